"""CPU checks: scene sizes, the oracle's block-CSR structure against a brute-force recomputation, and the C-ABI
library (loads, exports every declared symbol, reports errors with the reference's exception types)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from drape_amd import _ffi, scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_config_sizes_match_the_grid_formulas(oracle):
    """S = 2 (N-1)^2, H = 3 (N-1)^2 - 2 (N-1); block counts of the 32x32 and 256x256 sheets."""
    for n, nnzb in ((32, 12556), (64, None)):
        s = oracle.OracleSolver()
        s.build(scenes.c1_sheet(n=n))
        sz = s.structure_sizes()
        assert sz["n_particles"] == n * n
        assert sz["n_stretch"] == 2 * (n - 1) ** 2
        assert sz["n_hinges"] == 3 * (n - 1) ** 2 - 2 * (n - 1)
        if nnzb:
            assert sz["n_blocks"] == nnzb


def _brute_pattern(n, stretch_v, hinge_v, seam_ab):
    cols = [{i} for i in range(n)]
    for vs in list(stretch_v) + list(hinge_v) + list(seam_ab):
        for a in vs:
            cols[a].update(int(b) for b in vs)
    rs = [0]
    ci = []
    for i in range(n):
        ci.extend(sorted(cols[i]))
        rs.append(len(ci))
    return np.array(rs), np.array(ci)


def _seamed():
    import physics_cases as pc
    return scenes.Scene(garments=[pc.sewn_pair(gap=0.03, w=0.06, h=0.01)])


@pytest.mark.parametrize("make_scene", [lambda: scenes.c1_sheet(n=12), _seamed,
                                        lambda: scenes.c3_tube(nx=8, ny=6, spacing=0.02)])
def test_oracle_structure_matches_brute_force(oracle, make_scene):
    sc = make_scene()
    s = oracle.OracleSolver()
    s.build(sc)
    st, t = s.structure(), s.terms()
    rs, ci = _brute_pattern(sc.n, t["stretch_v"], t["hinge_v"], t["seam_ab"])
    np.testing.assert_array_equal(st["row_start"], rs)
    np.testing.assert_array_equal(st["col_index"], ci)
    assert np.all(ci[st["diag_slot"]] == np.arange(sc.n))

    def slot(r, c):
        row = ci[rs[r]:rs[r + 1]]
        k = np.searchsorted(row, c)
        assert row[k] == c
        return rs[r] + k
    for key, vkey, m in (("stretch_slots", "stretch_v", 3), ("hinge_slots", "hinge_v", 4), ("seam_slots", "seam_ab", 2)):
        for term, slots in zip(t[vkey], st[key]):
            expect = [slot(term[a], term[b]) for a in range(m) for b in range(m)]
            np.testing.assert_array_equal(slots, expect)
    # gather lists cover every contribution exactly once (gpu_solver_core.cpp:184-224)
    g = s.gather_lists()
    S, H, K = len(t["stretch_v"]), len(t["hinge_v"]), len(t["seam_ab"])
    assert len(g["grad_idx"]) == sc.n + 3 * S + 4 * H + 2 * K
    assert sorted(g["block_idx"]) == list(range(sc.n + 9 * S + 16 * H + 4 * K))


def test_tube_has_seams_and_seam_hinges(oracle):
    sc = scenes.c3_tube(nx=8, ny=6, spacing=0.02)
    s = oracle.OracleSolver()
    s.build(sc)
    sz = s.structure_sizes()
    assert sz["n_seams"] == 2 * 6
    interior = 2 * (3 * 7 * 5 - 7 - 5)  # per panel 3 (nx-1)(ny-1) - (nx-1) - (ny-1)
    assert sz["n_hinges"] == interior + 2 * (6 - 1)


def test_hinge_order_is_edge_key_order(oracle):
    """Hinges follow the (min, max) edge key order of the reference's std::map (sim_model.cpp:125-146)."""
    s = oracle.OracleSolver()
    s.build(scenes.c1_sheet(n=6))
    hv = s.terms()["hinge_v"]
    keys = [(min(h[2], h[3]), max(h[2], h[3])) for h in hv]
    assert keys == sorted(keys)
    assert all(h[2] < h[3] for h in hv)


def test_library_exports_every_declared_symbol():
    lib = _ffi.LIB_PATH
    assert os.path.exists(lib), "libdrape_hip.so not built"
    header = open(os.path.join(REPO, "include", "drape_hip.h")).read()
    declared = set(re.findall(r"\b(drape_[a-z0-9_]+)\s*\(", header))
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert declared and declared <= exported, declared - exported
    ctypes.CDLL(lib)  # loads without a GPU


def test_settings_default_is_the_reference_default():
    lib = _ffi.load_library()
    s = _ffi.Settings()
    assert lib.drape_settings_default(ctypes.byref(s)) == 0
    assert tuple(s.gravity) == (0.0, -9.81, 0.0)
    assert (s.collision_offset, s.assembly_closing_speed, s.assembly_extra_damping, s.damping_scale) == (
        0.003, 0.25, 10.0, 5.0)
    assert (s.strain_limit_stiffness_scale, s.seam_stiffness, s.friction_slip_speed) == (100.0, 1e4, 3e-4)
    assert (s.max_cg_iterations, s.cg_tolerance) == (50, 1e-2)
    assert scenes.DEFAULT_SETTINGS["gravity"] == (0.0, -9.81, 0.0)


def test_errors_map_to_reference_exception_types():
    lib = _ffi.load_library()
    assert lib.drape_settings_default(None) == _ffi.DRAPE_ERR_INVALID_ARGUMENT
    assert b"null" in lib.drape_last_error()
    # Without a gfx950 device the backend is unavailable and creation fails with a runtime error naming "hip"
    import drape_amd
    if drape_amd.hip_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert lib.drape_solver_create(ctypes.byref(h)) == _ffi.DRAPE_ERR_RUNTIME
    assert lib.drape_last_error().startswith(b"hip:")
    with pytest.raises(RuntimeError, match="hip"):
        drape_amd.HipSolver()


def test_scene_descriptor_rejects_arrays_shorter_than_the_mesh():
    """The C descriptor carries one vertex count and one triangle count per garment: a garment whose positions or
    per-vertex / per-triangle arrays disagree with them is refused before any pointer is handed over
    (sim_model.cpp:82-85: "one initial position per mesh vertex required")."""
    import copy
    sc = scenes.c1_sheet(n=8)
    _ffi.SceneHolder(sc)  # the scene as built packs
    cases = [("initial_positions", lambda g: g.initial_positions[:-1], "one initial position per mesh vertex"),
             ("vertex_area", lambda g: g.mesh.vertex_area[:-1], "vertex_area needs one entry per vertex"),
             ("triangle_piece", lambda g: g.mesh.triangle_piece[:-1], "triangle_piece needs one entry per triangle")]
    for field, cut, message in cases:
        bad = copy.deepcopy(sc)
        g = bad.garments[0]
        if field == "initial_positions":
            g.initial_positions = cut(g)
        else:
            setattr(g.mesh, field, cut(g))
        with pytest.raises(ValueError, match=message):
            _ffi.SceneHolder(bad)


def test_cpp_wrapper_header_compiles():
    """include/drape/sim/hip_solver.hpp (the C++ Solver-surface wrapper) compiles with the host compiler."""
    src = '#include "drape/sim/hip_solver.hpp"\nint main() { return drape::sim::hip::isAvailable() ? 0 : 1; }\n'
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(REPO, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
