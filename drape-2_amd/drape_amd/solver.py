"""Python mirror of the reference's Solver surface (libs/sim/include/drape/sim/solver.hpp:85-98) over the HIP
C-ABI (include/drape_hip.h).

`HipSolver` is the `SolverBackend::Hip` of `make_solver` (solver_factory.hpp:12-20): build / update_fabric /
set_phase / step / read_positions / stats / snapshot / restore / range / particle_count, with the reference's
exception types mapped to ValueError (invalid_argument), IndexError (out_of_range), RuntimeError
(runtime_error, naming "hip") and AssertionError (logic_error). It fails loudly when the library is not built
or no gfx950 device is present; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import (Counters, PcgReport, SceneDesc, SceneHolder, Stats, StructureSizes, load_library, ptr)


def _lib():
    return load_library()


def _check(code: int) -> None:
    _ffi.check(code, _lib().drape_last_error)


def hip_available() -> bool:
    """isAvailable(SolverBackend::Hip): the library loads and device 0 is a gfx950."""
    try:
        return bool(_lib().drape_hip_available())
    except (RuntimeError, OSError):
        return False


def default_settings() -> dict:
    s = _ffi.Settings()
    _check(_lib().drape_settings_default(C.byref(s)))
    out = {n: getattr(s, n) for n, _ in s._fields_ if n != "pad0"}
    out["gravity"] = tuple(s.gravity)
    return out


PERSISTENT = (1, 8, 2)  # the persistent solve's shape code


class HipSolver:
    """One MI355X solver instance (one segment per scene of a batched build)."""

    backend = "hip"

    def __init__(self, graph: bool = True, pcg_variant: int = 1, pcg_shape: tuple = (0, 0)):
        self._h = C.c_void_p()
        _check(_lib().drape_solver_create(C.byref(self._h)))
        self._n = 0
        if not graph:
            self.set_graph(False)
        if pcg_variant != 1:
            self.set_pcg_variant(pcg_variant)
        if tuple(pcg_shape)[:2] != (0, 0):
            self.set_pcg_shape(*pcg_shape)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _ffi._lib is not None:
            _ffi._lib.drape_solver_destroy(h)
            self._h = None

    # ---- Solver surface ----

    def build(self, scene) -> None:
        holder = SceneHolder(scene)
        _check(_lib().drape_solver_build(self._h, C.byref(holder.desc)))
        self._refresh_count()

    def build_batch(self, scenes: list) -> None:
        holders = [SceneHolder(s) for s in scenes]
        arr = (SceneDesc * max(1, len(holders)))(*[h.desc for h in holders])
        _check(_lib().drape_solver_build_batch(self._h, len(holders), arr))
        self._refresh_count()

    def _refresh_count(self):
        n = C.c_size_t()
        _check(_lib().drape_solver_particle_count(self._h, C.byref(n)))
        self._n = n.value

    def update_fabric(self, garment: str, fabrics: list) -> None:
        arr = (_ffi.Fabric * max(1, len(fabrics)))(*[_ffi.fabric_struct(f) for f in fabrics])
        _check(_lib().drape_solver_update_fabric(self._h, garment.encode(), len(fabrics), arr))

    def set_phase(self, garment: str, phase: int) -> None:
        _check(_lib().drape_solver_set_phase(self._h, garment.encode(), int(phase)))

    def step(self, dt=1.0 / 60.0, substeps=2, iterations=2) -> None:
        _check(_lib().drape_solver_step(self._h, float(dt), int(substeps), int(iterations)))

    def particle_count(self) -> int:
        return self._n

    def range(self, garment: str):
        off, cnt = C.c_uint32(), C.c_uint32()
        _check(_lib().drape_solver_range(self._h, garment.encode(), C.byref(off), C.byref(cnt)))
        return off.value, cnt.value

    def segment_count(self) -> int:
        n = C.c_uint32()
        _check(_lib().drape_solver_segment_count(self._h, C.byref(n)))
        return n.value

    def segment_range(self, k: int):
        off, cnt = C.c_uint32(), C.c_uint32()
        _check(_lib().drape_solver_segment_range(self._h, int(k), C.byref(off), C.byref(cnt)))
        return off.value, cnt.value

    def read_positions(self) -> np.ndarray:
        out = np.empty((self._n, 3), dtype=np.float32)
        _check(_lib().drape_solver_read_positions(self._h, ptr(out, C.c_float), self._n))
        return out

    def read_velocities(self) -> np.ndarray:
        out = np.empty((self._n, 3), dtype=np.float32)
        _check(_lib().drape_solver_read_velocities(self._h, ptr(out, C.c_float), self._n))
        return out

    def stats(self) -> dict:
        st = Stats()
        _check(_lib().drape_solver_stats(self._h, C.byref(st)))
        return st.as_dict()

    def read_positions_normals(self):
        """(positions, unit vertex normals) of the current state in particle order: every garment's triangles,
        seam pairs welded."""
        pos = np.empty((self._n, 3), np.float32)
        nrm = np.empty((self._n, 3), np.float32)
        _check(_lib().drape_solver_read_positions_normals(self._h, ptr(pos, C.c_float), ptr(nrm, C.c_float), self._n))
        return pos, nrm

    def segment_stats(self, first: int = 0, count: int | None = None) -> list:
        """Stats of each scene of a batched build (segments [first, first + count)), one device reduction."""
        if count is None:
            count = self.segment_count() - first
        arr = (Stats * max(count, 1))()
        _check(_lib().drape_solver_segment_stats(self._h, first, count, arr))
        return [arr[i].as_dict() for i in range(count)]

    def snapshot(self) -> bytes:
        size = C.c_size_t()
        _check(_lib().drape_solver_snapshot_size(self._h, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _check(_lib().drape_solver_snapshot_save(self._h, buf, size.value))
        return buf.raw

    def restore(self, blob: bytes) -> None:
        buf = C.create_string_buffer(blob, len(blob))
        _check(_lib().drape_solver_snapshot_restore(self._h, buf, len(blob)))

    def counters(self) -> dict:
        c = Counters()
        _check(_lib().drape_solver_counters(self._h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self) -> None:
        _check(_lib().drape_solver_reset_counters(self._h))

    def set_graph(self, enabled: bool) -> None:
        _check(_lib().drape_solver_set_graph(self._h, 1 if enabled else 0))

    def set_pcg_variant(self, variant: int) -> None:
        """1: one kernel per PCG iteration (default); 0: the two-kernel standard form."""
        _check(_lib().drape_solver_set_pcg_variant(self._h, int(variant)))

    def set_pcg_shape(self, waves: int, slots: int, halo: int = 0) -> None:
        """Launch shape of the one-kernel PCG iteration: (wavefronts per 64-row slice, slots per batch) in
        (1, 8), (1, 4), (2, 8), (4, 8), or (0, 0) to tune (default); halo=1 (with (1, 8) or (1, 4)) gathers from
        an LDS copy of each tile's column window, for matrices whose every tile spans at most 1,536 rows; halo=2
        (with (1, 8)) is the persistent solve: one launch per PCG solve, each scene's tiles resident together
        (PERSISTENT)."""
        _check(_lib().drape_solver_set_pcg_shape(self._h, int(waves), int(slots), int(halo)))

    def pcg_shape(self) -> tuple:
        """(waves, slots, halo) of the shape in use, halo 0, 1 or 2 (the persistent solve); (0, 0, 0) while still
        tuning."""
        w, k, h = C.c_int32(), C.c_int32(), C.c_int32()
        _check(_lib().drape_solver_pcg_shape(self._h, C.byref(w), C.byref(k), C.byref(h)))
        return (w.value, k.value, h.value)

    def set_solve_spin_cap(self, polls: int) -> None:
        """Polls of a persistent solve's barrier before a workgroup gives up (the solve is then re-run with one
        launch per iteration and the persistent form switched off)."""
        _check(_lib().drape_solver_set_solve_spin_cap(self._h, int(polls)))

    def persistent_available(self) -> bool:
        a = C.c_int32()
        _check(_lib().drape_solver_persistent_available(self._h, C.byref(a)))
        return bool(a.value)

    def set_spmv_sampling(self, every: int) -> None:
        _check(_lib().drape_solver_set_spmv_sampling(self._h, int(every)))

    def spmv_timing(self) -> dict:
        n, us, b = C.c_uint64(), C.c_double(), C.c_double()
        _check(_lib().drape_solver_spmv_timing(self._h, C.byref(n), C.byref(us), C.byref(b)))
        return {"launches": n.value, "mean_us": us.value, "bytes_per_launch": b.value}

    # ---- structure / kernel access ----

    def structure_sizes(self) -> dict:
        s = StructureSizes()
        _check(_lib().drape_solver_structure_sizes(self._h, C.byref(s)))
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
        _check(_lib().drape_solver_read_structure(self._h, *[ptr(out[k], C.c_uint32) for k in (
            "row_start", "col_index", "diag_slot", "stretch_slots", "hinge_slots", "seam_slots")]))
        return out

    def set_refinement(self, restarts: int) -> None:
        """FP64 true-residual continuations per PCG solve (0: off; default 4)."""
        _check(_lib().drape_solver_set_refinement(self._h, int(restarts)))

    def pcg_report(self) -> dict:
        """Why the PCG solves ended and their FP64 true relative residuals (drape_pcg_report)."""
        r = PcgReport()
        _check(_lib().drape_solver_pcg_report(self._h, C.byref(r)))
        return r.as_dict()

    def gather_lists(self) -> dict:
        """The assembly's summation order in the reference's gather-list numbering (gpu_solver_core.cpp:184-224)."""
        ng, nh = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        _check(_lib().drape_solver_gather_sizes(self._h, ptr(ng, C.c_uint32), ptr(nh, C.c_uint32)))
        sz = self.structure_sizes()
        out = {
            "grad_start": np.empty(sz["n_particles"] + 1, np.uint32),
            "grad_idx": np.empty(int(ng[0]), np.uint32),
            "block_start": np.empty(sz["n_blocks"] + 1, np.uint32),
            "block_idx": np.empty(int(nh[0]), np.uint32),
        }
        _check(_lib().drape_solver_read_gather(self._h, *[ptr(out[k], C.c_uint32) for k in (
            "grad_start", "grad_idx", "block_start", "block_idx")]))
        return out

    def terms(self) -> dict:
        sz = self.structure_sizes()
        out = {
            "stretch_v": np.empty((sz["n_stretch"], 3), np.uint32),
            "hinge_v": np.empty((sz["n_hinges"], 4), np.uint32),
            "hinge_k": np.empty(sz["n_hinges"], np.float64),
            "seam_ab": np.empty((sz["n_seams"], 2), np.uint32),
        }
        _check(_lib().drape_solver_read_terms(self._h, ptr(out["stretch_v"], C.c_uint32),
                                              ptr(out["hinge_v"], C.c_uint32), ptr(out["hinge_k"], C.c_double),
                                              ptr(out["seam_ab"], C.c_uint32)))
        return out

    def debug_assemble(self, h: float, u: np.ndarray):
        sz = self.structure_sizes()
        ua = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        e = C.c_double()
        grad = np.empty((self._n, 3), np.float32)
        blocks = np.empty((sz["n_blocks"], 3, 3), np.float32)
        _check(_lib().drape_solver_debug_assemble(self._h, float(h), ptr(ua, C.c_float), C.byref(e),
                                                  ptr(grad, C.c_float), ptr(blocks, C.c_float)))
        return e.value, grad, blocks

    # ---- learned warm start ----

    def set_warm_start(self, net) -> None:
        """Attach a `learn.WarmStartNet` (None restores the previous-velocity guess): each substep then seeds
        Newton at the inertial target plus h times the predicted velocity residual when that lowers the energy.
        Keep `net` alive."""
        self._warm_net = net
        _check(_lib().drape_solver_set_warm_start(self._h, net.handle if net is not None else None))

    def frame_inputs(self) -> dict:
        """Network inputs of the current state: x, v, dv_prev (the last step's realised dv), phi_normal."""
        out = {"x": np.empty((self._n, 3), np.float32), "v": np.empty((self._n, 3), np.float32),
               "dv_prev": np.empty((self._n, 3), np.float32), "phi_normal": np.empty((self._n, 4), np.float32)}
        _check(_lib().drape_solver_frame_inputs(self._h, ptr(out["x"], C.c_float), ptr(out["v"], C.c_float),
                                                ptr(out["dv_prev"], C.c_float), ptr(out["phi_normal"], C.c_float)))
        return out

    def mesh_graph(self):
        """The triangle-edge graph in particle order: (row_start, nbr)."""
        size = C.c_size_t()
        _check(_lib().drape_solver_mesh_graph_size(self._h, C.byref(size)))
        rs = np.empty(self._n + 1, np.uint32)
        nb = np.empty(max(size.value, 1), np.uint32)
        _check(_lib().drape_solver_mesh_graph(self._h, ptr(rs, C.c_uint32), ptr(nb, C.c_uint32)))
        return rs, nb[:size.value]

    def debug_system(self) -> dict:
        """The last Newton iteration's system: blocks (slot order), grad, block-Jacobi inverse, contact flags and
        contact normals (xyz, offset)."""
        sz = self.structure_sizes()
        out = {
            "blocks": np.empty((sz["n_blocks"], 3, 3), np.float32),
            "grad": np.empty((self._n, 3), np.float32),
            "precond": np.empty((self._n, 3, 3), np.float32),
            "contact_flags": np.empty(self._n, np.uint32),
            "contact_normals": np.empty((self._n, 4), np.float32),
        }
        _check(_lib().drape_solver_debug_system(self._h, ptr(out["blocks"], C.c_float), ptr(out["grad"], C.c_float),
                                                ptr(out["precond"], C.c_float),
                                                ptr(out["contact_flags"], C.c_uint32),
                                                ptr(out["contact_normals"], C.c_float)))
        return out

    def debug_spmv(self, x: np.ndarray) -> np.ndarray:
        xa = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        out = np.empty((self._n, 3), np.float32)
        _check(_lib().drape_solver_debug_spmv(self._h, ptr(xa, C.c_float), ptr(out, C.c_float)))
        return out


def make_solver(backend: str = "hip", **kw):
    """makeSolver (solver_factory.cpp:38-48) for this library's backend."""
    if backend != "hip":
        raise RuntimeError(f"{backend}: not built in this library (only 'hip')")
    return HipSolver(**kw)


# ---- standalone kernels ----

def body_meshes(components) -> tuple:
    """ctypes BodyMesh array for [(vertices [n, 3], triangles [t, 3]), ...] and the arrays it points into."""
    keep = []
    arr = (_ffi.BodyMesh * max(len(components), 1))()
    for i, (v, t) in enumerate(components):
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(t, np.uint32).reshape(-1, 3)
        keep += [v, t]
        arr[i] = _ffi.BodyMesh(ptr(v, C.c_double), len(v), ptr(t, C.c_uint32), len(t))
    return arr, keep


def bake_collision_field(components, cell_size=0.004, padding=0.05, band=0.02):
    """bakeCollisionField on the device (collision_bake.cpp:98-208): the signed distance field of the union of
    closed triangle meshes [(vertices, triangles), ...] as a scenes.Field."""
    from .scenes import Field
    arr, keep = body_meshes(components)
    opt = _ffi.BakeOptions(cell_size, padding, band)
    desc = _ffi.FieldDesc()
    _check(_lib().drape_bake_field_grid(arr, len(components), C.byref(opt), C.byref(desc)))
    nx, ny, nz = desc.dims
    values = np.empty((nz, ny, nx), np.float32)
    _check(_lib().drape_bake_collision_field(arr, len(components), C.byref(opt), ptr(values, C.c_float)))
    del keep
    return Field(origin=tuple(desc.origin), cell_size=desc.cell_size, dims=(nx, ny, nz), band=desc.band,
                 values=values)


def vertex_normals(positions, triangles, weld_pairs=None) -> np.ndarray:
    """Unit vertex normals on the device (vertexNormals, libs/render/src/normals.cpp:5-24): area-weighted sums of
    the triangles' cross products, weld pairs summed in order, normalised; (0, 1, 0) without area."""
    x = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    w = np.ascontiguousarray(weld_pairs if weld_pairs is not None else np.zeros((0, 2)), np.uint32).reshape(-1, 2)
    out = np.empty_like(x)
    _check(_lib().drape_vertex_normals(ptr(x, C.c_float), len(x), ptr(t, C.c_uint32), len(t), ptr(w, C.c_uint32),
                                       len(w), ptr(out, C.c_float)))
    return out


def bsr_spmv(row_start, col_index, blocks, pinned, x, reps: int = 1):
    """y = A x on the device with the solver's SELL-64 SpMV kernel; returns (y, mean kernel microseconds)."""
    n = len(row_start) - 1
    rs = np.ascontiguousarray(row_start, np.uint32)
    ci = np.ascontiguousarray(col_index, np.uint32)
    bl = np.ascontiguousarray(blocks, np.float32).reshape(-1)
    pn = None if pinned is None else np.ascontiguousarray(pinned, np.uint8)
    xa = np.ascontiguousarray(x, np.float32).reshape(-1)
    y = np.empty((n, 3), np.float32)
    us = C.c_double()
    _check(_lib().drape_bsr_spmv(n, len(ci), ptr(rs, C.c_uint32), ptr(ci, C.c_uint32), ptr(bl, C.c_float),
                                 ptr(pn, C.c_uint8), ptr(xa, C.c_float), ptr(y, C.c_float), int(reps), C.byref(us)))
    return y, us.value


def eval_terms(kind: int, x: np.ndarray, params: np.ndarray, fabrics=None, limit_scale: float = 100.0):
    """FP32 device term evaluation: kind 0 stretch, 1 hinge, 2 seam, 3 friction (see drape_eval_terms)."""
    N = {0: 3, 1: 4, 2: 2, 3: 1}[kind]
    xa = np.ascontiguousarray(x, np.float32).reshape(-1)
    pa = np.ascontiguousarray(params, np.float32).reshape(-1)
    n = len(xa) // (3 * N)
    fab = None
    if kind == 0:
        fab = (_ffi.Fabric * n)(*[_ffi.fabric_struct(f) for f in fabrics])
    e = np.empty(n, np.float32)
    g = np.empty((n, N, 3), np.float32)
    h = np.empty((n, N, N, 3, 3), np.float32)
    _check(_lib().drape_eval_terms(int(kind), n, ptr(xa, C.c_float), ptr(pa, C.c_float),
                                   fab if fab is not None else C.cast(None, C.POINTER(_ffi.Fabric)),
                                   float(limit_scale), ptr(e, C.c_float), ptr(g, C.c_float), ptr(h, C.c_float)))
    return e, g, h


def field_query(field, points: np.ndarray):
    keep: list = []
    fd = _ffi.field_struct(field, keep)
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    phi = np.empty(len(p), np.float32)
    nrm = np.empty((len(p), 3), np.float32)
    _check(_lib().drape_field_query(C.byref(fd), len(p), ptr(p, C.c_float), ptr(phi, C.c_float), ptr(nrm, C.c_float)))
    return phi, nrm
