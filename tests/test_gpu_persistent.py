"""The persistent PCG solve: one launch per solve, each tile's workgroup keeping its slices' matrix and private vectors
in registers, the tiles of each scene meeting at a barrier of their own every iteration.

It computes every iteration as the launch-per-iteration form with the 1x8 halo shape does, so the two must agree bit
for bit: on the 32x32 sheet, through the 65,536-particle drop's first contact, and segment by segment on a batch of
uneven sheets (whose scenes never wait for each other). Every wait is bounded: a barrier that gives up sets the
abort flag, every workgroup leaves, and the solver re-runs that solve with one launch per iteration.
"""
import numpy as np
import pytest

from drape_amd import scenes

pytestmark = pytest.mark.gpu

PERSISTENT = (1, 8, 2)
HALO_LAUNCHES = (1, 8, 1)


def _run(hip, sc, shape, steps, spin_cap=None, batch=False):
    g = hip.HipSolver(pcg_shape=shape)
    if spin_cap is not None:
        g.set_solve_spin_cap(spin_cap)
    if batch:
        g.build_batch(sc)
    else:
        g.build(sc)
    for _ in range(steps):
        g.step(1 / 60, 1, 5)
    return g


def _same(a, b):
    np.testing.assert_array_equal(a.read_positions(), b.read_positions())
    np.testing.assert_array_equal(a.read_velocities(), b.read_velocities())
    assert a.counters() == b.counters()
    ra, rb = a.pcg_report(), b.pcg_report()
    for key in ("solves", "exit_tolerance", "exit_cap", "exit_curvature", "restarts"):
        assert ra[key] == rb[key], key
    assert ra["true_rel_max"] == rb["true_rel_max"]


def test_persistent_solve_equals_launches_on_the_small_sheet(hip):
    sc = scenes.c1_sheet()
    p = _run(hip, sc, PERSISTENT, 12)
    assert p.persistent_available() and p.pcg_shape() == PERSISTENT
    assert p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, sc, HALO_LAUNCHES, 12)
    print(f"32x32 sheet, 12 steps: counters {p.counters()}", flush=True)
    _same(p, launches)


def test_persistent_solve_equals_launches_through_first_contact_at_65k(hip):
    """256 tiles, one per CU: the largest single scene the persistent form takes on MI355X. 26 steps reach the
    sphere (first contact near step 23), where the solves are longest and the refinement continues some of them."""
    sc = scenes.c2_sheet()
    p = _run(hip, sc, PERSISTENT, 26)
    assert p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, sc, HALO_LAUNCHES, 26)
    print(f"65,536-particle drop, 26 steps: counters {p.counters()}, report {p.pcg_report()}", flush=True)
    _same(p, launches)


def test_persistent_batch_segments_equal_their_launch_runs(hip):
    """Eight uneven sheets (3,312 to 10,856 particles, 197 tiles) in one persistent launch per solve: each
    segment's barrier counts only its own tiles, and the batch equals the launch form segment for segment; the
    smallest and largest segments also equal their single-scene persistent runs."""
    batch = scenes.c4_batch(count=8, lo=2000, hi=12000, tolerance=1e-6, max_cg=2000)
    steps = 6
    p = _run(hip, batch, PERSISTENT, steps, batch=True)
    assert p.persistent_available() and p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, batch, HALO_LAUNCHES, steps, batch=True)
    print(f"8-sheet batch, {steps} steps: counters {p.counters()}", flush=True)
    _same(p, launches)
    xb = p.read_positions()
    sizes = [sc.n for sc in batch]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        off, cnt = p.segment_range(k)
        single = _run(hip, batch[k], PERSISTENT, steps)
        np.testing.assert_array_equal(xb[off:off + cnt], single.read_positions())


def test_persistent_queue_on_the_full_64_sheet_batch(hip):
    """The 64 uneven sheets (891,781 particles, 3,517 tiles) outnumber the CUs, so the persistent solve queues them:
    each sheet's tiles (at most 153) resident together, the sheets passing through the CUs largest expected work
    first. Three steps through the drop equal the launch form bit for bit, with no barrier giving up."""
    batch = scenes.c4_batch()
    steps = 3
    p = _run(hip, batch, PERSISTENT, steps, batch=True)
    rep = p.pcg_report()
    print(f"64-sheet batch, {steps} steps: counters {p.counters()}, aborts {rep['persistent_aborts']}", flush=True)
    assert p.persistent_available() and rep["persistent_aborts"] == 0
    launches = _run(hip, batch, HALO_LAUNCHES, steps, batch=True)
    _same(p, launches)


def test_persistent_solve_gives_up_and_the_solve_is_redone_with_launches(hip):
    """A spin cap of zero makes every barrier wait of a multi-tile scene give up at once (forced from the host; the
    kernel still exits normally): the solver re-runs the solve with one launch per iteration, switches the
    persistent form off, and its results are the launch form's."""
    sc = scenes.c2_sheet(n=96)  # 36 tiles: three barrier groups
    p = _run(hip, sc, PERSISTENT, 4, spin_cap=0)
    rep = p.pcg_report()
    print(f"spin cap 0: report {rep}, shape now {p.pcg_shape()}", flush=True)
    assert rep["persistent_aborts"] == 1
    assert not p.persistent_available()
    assert p.pcg_shape() != PERSISTENT
    launches = _run(hip, sc, HALO_LAUNCHES, 4)
    np.testing.assert_array_equal(p.read_positions(), launches.read_positions())
    assert p.counters()["newton_iterations"] == launches.counters()["newton_iterations"]


def test_persistent_solve_two_tiles_per_cu_on_one_large_scene(hip):
    """A 288 x 288 sheet is one scene of 324 tiles: more than the CUs, so its tiles run two to a CU, all resident
    together in one launch. Three steps equal the launch form bit for bit."""
    sc = scenes.c2_sheet(n=288)
    p = _run(hip, sc, PERSISTENT, 3)
    assert p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, sc, HALO_LAUNCHES, 3)
    _same(p, launches)


def test_persistent_solve_on_the_seamed_tube(hip):
    """The 120,000-particle seamed tube: 469 tiles, run by pairs (235 workgroups, one per CU; the last pair has one
    tile), each pair's columns in up to eight row ranges (its panels' neighbourhoods and the seam partners' on the
    other panel), seam rows with 9 upper and 9 lower slots (the slots past the eighth held in the pair's LDS list).
    Three steps equal one launch per iteration bit for bit."""
    sc = scenes.c3_tube()
    p = _run(hip, sc, PERSISTENT, 3)
    print(f"tube, 3 steps: counters {p.counters()}, aborts {p.pcg_report()['persistent_aborts']}", flush=True)
    assert p.persistent_available() and p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, sc, (1, 8, 0), 3)
    _same(p, launches)


def test_persistent_pairs_queued_on_a_batch_of_seamed_tubes(hip):
    """Three seamed tubes of 120,000, 78,000 and 48,000 particles batched: 962 tiles, 482 pairs, more than the CUs,
    so the seam rows' pairs are queued through the CUs tube by tube (each tube's pairs resident together, at most
    235). Two steps equal one launch per iteration bit for bit, tube by tube, with no barrier giving up."""
    batch = [scenes.c3_tube(), scenes.c3_tube(nx=150, ny=260, seed=5), scenes.c3_tube(nx=120, ny=200, seed=6)]
    steps = 2
    p = _run(hip, batch, PERSISTENT, steps, batch=True)
    rep = p.pcg_report()
    print(f"3-tube batch, {steps} steps: counters {p.counters()}, aborts {rep['persistent_aborts']}", flush=True)
    assert p.persistent_available() and p.pcg_shape() == PERSISTENT and rep["persistent_aborts"] == 0
    launches = _run(hip, batch, (1, 8, 0), steps, batch=True)
    _same(p, launches)


def _sheet_with_long_rows():
    """The 32 x 32 sheet (pinned at vertices 0 and 31) with extra seam springs that give a few rows 9 to 12 lower or
    upper slots: row 500 is sewn to the pinned corner 0, to 40-44 and to 900-904, row 600 to 0 and 30-33 (every pair
    a seam of its own, so no seam hinges)."""
    sc = scenes.c1_sheet()
    m = sc.garments[0].mesh
    pairs = [(500, b) for b in (0, 40, 41, 42, 43, 44, 900, 901, 902, 903, 904)] + [(600, b) for b in (0, 30, 31, 32, 33)]
    m.seam_pairs = np.array([(a, b, k, 0) for k, (a, b) in enumerate(pairs)], dtype=np.uint32)
    return sc


def test_persistent_solve_with_rows_past_sixteen_slots_in_all(hip):
    """Rows with up to four lower and four upper slots past the eighth, among them a pinned column (a real slot whose
    gathered value is zero), and lanes of the same slices with none: the LDS list of the extra slots and the padding
    slots' zero products keep the persistent solve bit-identical to one launch per iteration."""
    sc = _sheet_with_long_rows()
    probe = hip.HipSolver()
    probe.build(sc)
    st = probe.structure()
    rs, ci = st["row_start"].astype(np.int64), st["col_index"].astype(np.int64)
    lower = np.array([np.searchsorted(ci[rs[r]:rs[r + 1]], r) for r in range(len(rs) - 1)])
    upper = np.diff(rs) - lower
    print(f"largest lower / upper slot counts: {lower.max()} / {upper.max()}", flush=True)
    assert lower.max() > 9 and upper.max() > 9 and lower.max() <= 16 and upper.max() <= 16
    p = _run(hip, sc, PERSISTENT, 4)
    assert p.persistent_available() and p.pcg_report()["persistent_aborts"] == 0
    launches = _run(hip, sc, (1, 8, 0), 4)
    assert np.all(np.isfinite(p.read_positions()))
    _same(p, launches)


def _scattered_sheet(n=64, seed=5):
    """A sheet whose vertices are numbered in random order: every tile's columns scatter over the whole matrix."""
    sc = scenes.c2_sheet(n=n, spacing=0.01)
    g = sc.garments[0]
    m = g.mesh
    perm = np.random.default_rng(seed).permutation(m.n)  # new index of vertex i is inv[i]
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.n)
    m.rest = m.rest[perm]
    m.vertex_area = m.vertex_area[perm]
    m.boundary = m.boundary[perm]
    m.vertex_piece = m.vertex_piece[perm]
    m.triangles = inv[m.triangles].astype(np.uint32)
    g.initial_positions = g.initial_positions[perm]
    return sc


def test_persistent_solve_is_refused_where_it_does_not_fit(hip):
    """A sheet numbered in random order has tiles whose columns need more than four row ranges or 1,536 rows, and a
    384 x 384 sheet is one scene of 576 tiles, more than the device holds at two a CU (a scene's tiles wait for each
    other, so they must all be resident): the persistent form is refused before anything is launched, and a tuned
    solver never picks it there."""
    for sc in (_scattered_sheet(), scenes.c2_sheet(n=384)):
        g = hip.HipSolver(pcg_shape=PERSISTENT)
        g.build(sc)
        assert not g.persistent_available()
        with pytest.raises(ValueError, match="persistent"):
            g.step(1 / 60, 1, 5)
