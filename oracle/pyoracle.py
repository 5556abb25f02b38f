"""ctypes binding of the FP64 CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker or the timed CPU baseline. `OracleSolver` exposes the same Python surface as the product's
`drape_amd.HipSolver`, so a test can run one scene on both backends (the reference's backend-parameterised
tests, libs/sim/tests/backends.hpp:14-47).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "drape-2_amd"))

from drape_amd import _ffi  # noqa: E402
from drape_amd._ffi import (BakeOptions, BodyMesh, Counters, FieldDesc, SceneDesc, SceneHolder, Stats, StructureSizes, f32p, f64p,  # noqa: E402
                            ptr, u8p, u32p)

LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build() -> None:
    """Compiles the oracle with its Makefile (g++, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    S = C.c_void_p
    sig = {
        "oracle_last_error": (C.c_char_p, []),
        "oracle_set_threads": (C.c_int, [C.c_int32]),
        "oracle_solver_create": (C.c_int, [C.POINTER(S)]),
        "oracle_solver_destroy": (None, [S]),
        "oracle_solver_build": (C.c_int, [S, C.POINTER(SceneDesc)]),
        "oracle_solver_update_fabric": (C.c_int, [S, C.c_char_p, C.c_uint32, C.POINTER(_ffi.Fabric)]),
        "oracle_solver_set_phase": (C.c_int, [S, C.c_char_p, C.c_int32]),
        "oracle_solver_step": (C.c_int, [S, C.c_double, C.c_int32, C.c_int32]),
        "oracle_solver_particle_count": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "oracle_solver_range": (C.c_int, [S, C.c_char_p, u32p, u32p]),
        "oracle_solver_read_positions": (C.c_int, [S, f32p, C.c_size_t]),
        "oracle_solver_read_velocities": (C.c_int, [S, f32p, C.c_size_t]),
        "oracle_solver_write_state": (C.c_int, [S, f32p, f32p, C.c_size_t]),
        "oracle_solver_stats": (C.c_int, [S, C.POINTER(Stats)]),
        "oracle_vertex_normals": (C.c_int, [f32p, C.c_uint32, u32p, C.c_uint32, u32p, C.c_uint32, f32p]),
        "oracle_bake_field_grid": (C.c_int, [C.POINTER(BodyMesh), C.c_uint32, C.POINTER(BakeOptions),
                                             C.POINTER(FieldDesc)]),
        "oracle_bake_collision_field": (C.c_int, [C.POINTER(BodyMesh), C.c_uint32, C.POINTER(BakeOptions), f32p]),
        "oracle_solver_snapshot_size": (C.c_int, [S, C.POINTER(C.c_size_t)]),
        "oracle_solver_snapshot_save": (C.c_int, [S, C.c_void_p, C.c_size_t]),
        "oracle_solver_snapshot_restore": (C.c_int, [S, C.c_void_p, C.c_size_t]),
        "oracle_solver_counters": (C.c_int, [S, C.POINTER(Counters)]),
        "oracle_solver_structure_sizes": (C.c_int, [S, C.POINTER(StructureSizes)]),
        "oracle_solver_read_structure": (C.c_int, [S, u32p, u32p, u32p, u32p, u32p, u32p]),
        "oracle_solver_read_terms": (C.c_int, [S, u32p, u32p, f64p, u32p]),
        "oracle_solver_gather_sizes": (C.c_int, [S, u32p, u32p]),
        "oracle_solver_read_gather": (C.c_int, [S, u32p, u32p, u32p, u32p]),
        "oracle_solver_debug_assemble": (C.c_int, [S, C.c_double, f64p, f64p, f64p, f64p]),
        "oracle_solver_debug_spmv": (C.c_int, [S, f64p, f64p]),
        "oracle_eval_stretch": (C.c_int, [f64p, f64p, C.c_double, C.POINTER(_ffi.Fabric), C.c_double, f64p, f64p,
                                          f64p]),
        "oracle_eval_hinge": (C.c_int, [f64p, C.c_double, C.c_double, f64p, f64p, f64p]),
        "oracle_eval_seam": (C.c_int, [f64p, C.c_double, C.c_double, f64p, f64p, f64p]),
        "oracle_eval_friction": (C.c_int, [f64p, f64p, f64p, C.c_double, C.c_double, f64p, f64p, f64p]),
        "oracle_strain": (C.c_int, [f64p, f64p, C.c_int32, f64p, f64p]),
        "oracle_dihedral": (C.c_int, [f64p, f64p, f64p]),
        "oracle_bending_coefficients": (C.c_int, [C.c_double, C.c_double, f64p, f64p]),
        "oracle_edge_bend_stiffness": (C.c_double, [C.c_double, C.c_double, f64p, f64p]),
        "oracle_field_query": (C.c_int, [C.POINTER(FieldDesc), C.c_uint32, f64p, f64p, f64p]),
        "oracle_collide_body": (C.c_int, [C.POINTER(FieldDesc), C.c_float, C.c_float, f32p, f32p]),
        "oracle_land_on_body": (C.c_int, [C.POINTER(FieldDesc), C.c_float, f32p, f32p, f32p]),
        "oracle_invert_spd3": (C.c_int, [f64p, f64p]),
        "oracle_bsr_pcg": (C.c_int, [C.c_uint32, u32p, u32p, f64p, u8p, f64p, C.c_double, C.c_int32, f64p,
                                     C.POINTER(C.c_int32)]),
        "oracle_bsr_pcg_contact": (C.c_int, [C.c_uint32, u32p, u32p, f64p, u8p, u8p, f64p, f64p, C.c_double,
                                             C.c_int32, f64p, C.POINTER(C.c_int32), f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int) -> None:
    _ffi.check(code, lib().oracle_last_error)


def set_threads(n: int) -> None:
    check(lib().oracle_set_threads(int(n)))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleSolver:
    """FP64 CpuSolver restatement behind the same Python surface as drape_amd.HipSolver."""

    backend = "cpu"

    def __init__(self):
        self._h = C.c_void_p()
        check(lib().oracle_solver_create(C.byref(self._h)))
        self._n = 0

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.oracle_solver_destroy(self._h)
            self._h = None

    def build(self, scene) -> None:
        holder = SceneHolder(scene)
        check(lib().oracle_solver_build(self._h, C.byref(holder.desc)))
        n = C.c_size_t()
        check(lib().oracle_solver_particle_count(self._h, C.byref(n)))
        self._n = n.value

    def update_fabric(self, garment: str, fabrics: list) -> None:
        arr = (_ffi.Fabric * max(1, len(fabrics)))(*[_ffi.fabric_struct(f) for f in fabrics])
        check(lib().oracle_solver_update_fabric(self._h, garment.encode(), len(fabrics), arr))

    def set_phase(self, garment: str, phase: int) -> None:
        check(lib().oracle_solver_set_phase(self._h, garment.encode(), int(phase)))

    def step(self, dt=1.0 / 60.0, substeps=2, iterations=2) -> None:
        check(lib().oracle_solver_step(self._h, float(dt), int(substeps), int(iterations)))

    def particle_count(self) -> int:
        return self._n

    def range(self, garment: str):
        off, cnt = C.c_uint32(), C.c_uint32()
        check(lib().oracle_solver_range(self._h, garment.encode(), C.byref(off), C.byref(cnt)))
        return off.value, cnt.value

    def read_positions(self) -> np.ndarray:
        out = np.empty((self._n, 3), dtype=np.float32)
        check(lib().oracle_solver_read_positions(self._h, ptr(out, C.c_float), self._n))
        return out

    def read_velocities(self) -> np.ndarray:
        out = np.empty((self._n, 3), dtype=np.float32)
        check(lib().oracle_solver_read_velocities(self._h, ptr(out, C.c_float), self._n))
        return out

    def write_state(self, x=None, v=None) -> None:
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float32)
        va = None if v is None else np.ascontiguousarray(v, dtype=np.float32)
        check(lib().oracle_solver_write_state(self._h, ptr(xa, C.c_float), ptr(va, C.c_float), self._n))

    def stats(self) -> dict:
        st = Stats()
        check(lib().oracle_solver_stats(self._h, C.byref(st)))
        return st.as_dict()

    def snapshot(self) -> bytes:
        size = C.c_size_t()
        check(lib().oracle_solver_snapshot_size(self._h, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        check(lib().oracle_solver_snapshot_save(self._h, buf, size.value))
        return buf.raw

    def restore(self, blob: bytes) -> None:
        buf = C.create_string_buffer(blob, len(blob))
        check(lib().oracle_solver_snapshot_restore(self._h, buf, len(blob)))

    def counters(self) -> dict:
        c = Counters()
        check(lib().oracle_solver_counters(self._h, C.byref(c)))
        return c.as_dict()

    # ---- structure / assembly access ----

    def structure_sizes(self) -> dict:
        s = StructureSizes()
        check(lib().oracle_solver_structure_sizes(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in s._fields_}

    def structure(self) -> dict:
        sz = self.structure_sizes()
        out = {
            "row_start": np.empty(sz["n_particles"] + 1, np.uint32),
            "col_index": np.empty(sz["n_blocks"], np.uint32),
            "diag_slot": np.empty(sz["n_particles"], np.uint32),
            "stretch_slots": np.empty((sz["n_stretch"], 9), np.uint32),
            "hinge_slots": np.empty((sz["n_hinges"], 16), np.uint32),
            "seam_slots": np.empty((sz["n_seams"], 4), np.uint32),
        }
        check(lib().oracle_solver_read_structure(self._h, *[ptr(out[k], C.c_uint32) for k in (
            "row_start", "col_index", "diag_slot", "stretch_slots", "hinge_slots", "seam_slots")]))
        return out

    def terms(self) -> dict:
        sz = self.structure_sizes()
        out = {
            "stretch_v": np.empty((sz["n_stretch"], 3), np.uint32),
            "hinge_v": np.empty((sz["n_hinges"], 4), np.uint32),
            "hinge_k": np.empty(sz["n_hinges"], np.float64),
            "seam_ab": np.empty((sz["n_seams"], 2), np.uint32),
        }
        check(lib().oracle_solver_read_terms(self._h, ptr(out["stretch_v"], C.c_uint32),
                                             ptr(out["hinge_v"], C.c_uint32), ptr(out["hinge_k"], C.c_double),
                                             ptr(out["seam_ab"], C.c_uint32)))
        return out

    def gather_lists(self) -> dict:
        ng, nh = C.c_uint32(), C.c_uint32()
        check(lib().oracle_solver_gather_sizes(self._h, C.byref(ng), C.byref(nh)))
        sz = self.structure_sizes()
        out = {
            "grad_start": np.empty(sz["n_particles"] + 1, np.uint32),
            "grad_idx": np.empty(ng.value, np.uint32),
            "block_start": np.empty(sz["n_blocks"] + 1, np.uint32),
            "block_idx": np.empty(nh.value, np.uint32),
        }
        check(lib().oracle_solver_read_gather(self._h, *[ptr(out[k], C.c_uint32) for k in (
            "grad_start", "grad_idx", "block_start", "block_idx")]))
        return out

    def debug_assemble(self, h: float, u: np.ndarray):
        sz = self.structure_sizes()
        ua = _d(u).reshape(-1)
        e = C.c_double()
        grad = np.empty((self._n, 3))
        blocks = np.empty((sz["n_blocks"], 3, 3))
        check(lib().oracle_solver_debug_assemble(self._h, float(h), ptr(ua, C.c_double), C.byref(e),
                                                 ptr(grad, C.c_double), ptr(blocks, C.c_double)))
        return e.value, grad, blocks

    def debug_spmv(self, x: np.ndarray) -> np.ndarray:
        xa = _d(x).reshape(-1)
        out = np.empty((self._n, 3))
        check(lib().oracle_solver_debug_spmv(self._h, ptr(xa, C.c_double), ptr(out, C.c_double)))
        return out


# ---- term math ----

def eval_stretch(x, dm_inv, area, fab, limit_scale=100.0):
    xa, d = _d(x).reshape(-1), _d(dm_inv).reshape(-1)
    e = C.c_double()
    g = np.empty((3, 3))
    h = np.empty((3, 3, 3, 3))
    f = _ffi.fabric_struct(fab)
    check(lib().oracle_eval_stretch(ptr(xa, C.c_double), ptr(d, C.c_double), float(area), C.byref(f),
                                    float(limit_scale), C.byref(e), ptr(g, C.c_double), ptr(h, C.c_double)))
    return e.value, g, h


def eval_hinge(x, K, rest_angle=0.0):
    xa = _d(x).reshape(-1)
    e = C.c_double()
    g = np.empty((4, 3))
    h = np.empty((4, 4, 3, 3))
    check(lib().oracle_eval_hinge(ptr(xa, C.c_double), float(K), float(rest_angle), C.byref(e), ptr(g, C.c_double),
                                  ptr(h, C.c_double)))
    return e.value, g, h


def eval_seam(x, target, K):
    xa = _d(x).reshape(-1)
    e = C.c_double()
    g = np.empty((2, 3))
    h = np.empty((2, 2, 3, 3))
    check(lib().oracle_eval_seam(ptr(xa, C.c_double), float(target), float(K), C.byref(e), ptr(g, C.c_double),
                                 ptr(h, C.c_double)))
    return e.value, g, h


def eval_friction(x, x0, n, mu_lambda, eps):
    xa, x0a, na = _d(x), _d(x0), _d(n)
    e = C.c_double()
    g = np.empty((1, 3))
    h = np.empty((1, 1, 3, 3))
    check(lib().oracle_eval_friction(ptr(xa, C.c_double), ptr(x0a, C.c_double), ptr(na, C.c_double),
                                     float(mu_lambda), float(eps), C.byref(e), ptr(g, C.c_double),
                                     ptr(h, C.c_double)))
    return e.value, g, h


def strain(x, dm_inv, which):
    xa, d = _d(x).reshape(-1), _d(dm_inv).reshape(-1)
    c = C.c_double()
    g = np.empty((3, 3))
    check(lib().oracle_strain(ptr(xa, C.c_double), ptr(d, C.c_double), int(which), C.byref(c), ptr(g, C.c_double)))
    return c.value, g


def dihedral(x):
    xa = _d(x).reshape(-1)
    t = C.c_double()
    g = np.empty((4, 3))
    check(lib().oracle_dihedral(ptr(xa, C.c_double), C.byref(t), ptr(g, C.c_double)))
    return t.value, g


def bending_coefficients(bend_warp, bend_weft):
    p, q = C.c_double(), C.c_double()
    check(lib().oracle_bending_coefficients(float(bend_warp), float(bend_weft), C.byref(p), C.byref(q)))
    return p.value, q.value


def edge_bend_stiffness(p, q, edge, grain):
    e, g = _d(edge), _d(grain)
    return lib().oracle_edge_bend_stiffness(float(p), float(q), ptr(e, C.c_double), ptr(g, C.c_double))


def field_query(field, points):
    keep: list = []
    fd = _ffi.field_struct(field, keep)
    p = _d(points).reshape(-1, 3)
    phi = np.empty(len(p))
    nrm = np.empty((len(p), 3))
    check(lib().oracle_field_query(C.byref(fd), len(p), ptr(p, C.c_double), ptr(phi, C.c_double),
                                   ptr(nrm, C.c_double)))
    return phi, nrm


def collide_body(field, radius, friction, x_prev, x):
    keep: list = []
    fd = _ffi.field_struct(field, keep)
    xp = np.ascontiguousarray(x_prev, dtype=np.float32)
    xa = np.array(x, dtype=np.float32)
    check(lib().oracle_collide_body(C.byref(fd), float(radius), float(friction), ptr(xp, C.c_float),
                                    ptr(xa, C.c_float)))
    return xa


def land_on_body(field, radius, frm, to):
    keep: list = []
    fd = _ffi.field_struct(field, keep)
    a = np.ascontiguousarray(frm, dtype=np.float32)
    b = np.ascontiguousarray(to, dtype=np.float32)
    out = np.empty(3, dtype=np.float32)
    check(lib().oracle_land_on_body(C.byref(fd), float(radius), ptr(a, C.c_float), ptr(b, C.c_float),
                                    ptr(out, C.c_float)))
    return out


def invert_spd3(a):
    aa = _d(a).reshape(-1)
    out = np.zeros(9)
    ok = lib().oracle_invert_spd3(ptr(aa, C.c_double), ptr(out, C.c_double))
    return bool(ok), out.reshape(3, 3)


def bsr_pcg(row_start, col_index, blocks, pinned, rhs, tol, max_iter):
    n = len(row_start) - 1
    rs = np.ascontiguousarray(row_start, np.uint32)
    ci = np.ascontiguousarray(col_index, np.uint32)
    bl = _d(blocks).reshape(-1)
    pn = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
    b = _d(rhs).reshape(-1)
    x = np.empty((n, 3))
    it = C.c_int32()
    check(lib().oracle_bsr_pcg(n, ptr(rs, C.c_uint32), ptr(ci, C.c_uint32), ptr(bl, C.c_double),
                               ptr(pn, C.c_uint8), ptr(b, C.c_double), float(tol), int(max_iter),
                               ptr(x, C.c_double), C.byref(it)))
    return x, it.value


def bake_collision_field(components, cell_size=0.004, padding=0.05, band=0.02):
    """bakeCollisionField (collision_bake.cpp:98-208), sequential FP64: a scenes.Field."""
    from drape_amd.scenes import Field
    from drape_amd.solver import body_meshes
    arr, keep = body_meshes(components)
    opt = BakeOptions(cell_size, padding, band)
    desc = FieldDesc()
    check(lib().oracle_bake_field_grid(arr, len(components), C.byref(opt), C.byref(desc)))
    nx, ny, nz = desc.dims
    values = np.empty((nz, ny, nx), np.float32)
    check(lib().oracle_bake_collision_field(arr, len(components), C.byref(opt), ptr(values, C.c_float)))
    del keep
    return Field(origin=tuple(desc.origin), cell_size=desc.cell_size, dims=(nx, ny, nz), band=desc.band,
                 values=values)


def vertex_normals(positions, triangles, weld_pairs=None) -> np.ndarray:
    """vertexNormals (normals.cpp:5-24) in FP32, sequential: unit normal per vertex."""
    x = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    w = np.ascontiguousarray(weld_pairs if weld_pairs is not None else np.zeros((0, 2)), np.uint32).reshape(-1, 2)
    out = np.empty_like(x)
    check(lib().oracle_vertex_normals(ptr(x, C.c_float), len(x), ptr(t, C.c_uint32), len(t), ptr(w, C.c_uint32),
                                      len(w), ptr(out, C.c_float)))
    return out


def bsr_pcg_contact(row_start, col_index, blocks, pinned, active, normals, rhs, tol, max_iter):
    """solveLinear with held contacts: (step, PCG iterations, contact force per row)."""
    n = len(row_start) - 1
    rs = np.ascontiguousarray(row_start, np.uint32)
    ci = np.ascontiguousarray(col_index, np.uint32)
    bl = _d(blocks).reshape(-1)
    pn = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
    ac = np.ascontiguousarray(active, np.uint8)
    nm = _d(normals).reshape(-1)
    b = _d(rhs).reshape(-1)
    x = np.empty((n, 3))
    force = np.empty(n)
    it = C.c_int32()
    check(lib().oracle_bsr_pcg_contact(n, ptr(rs, C.c_uint32), ptr(ci, C.c_uint32), ptr(bl, C.c_double),
                                       ptr(pn, C.c_uint8), ptr(ac, C.c_uint8), ptr(nm, C.c_double),
                                       ptr(b, C.c_double), float(tol), int(max_iter), ptr(x, C.c_double),
                                       C.byref(it), ptr(force, C.c_double)))
    return x, it.value, force
