"""CPU checks of the warm-start network's reference (tests/mpnn_ref.py): its backward pass against central
finite differences of its own FP64 loss, and the scene graph against a brute-force edge list."""
import numpy as np
import pytest

import mpnn_ref
from drape_amd import learn, scenes


def _small_case(seed=3, n=6):
    sc = scenes.hanging_sheet(size=0.05, spacing=0.01)
    rs, nbr, rest2 = learn.scene_graph(sc)
    m = len(rs) - 1
    rng = np.random.default_rng(seed)
    frame = {
        "x": np.asarray(sc.garments[0].initial_positions) + rng.normal(0, 2e-3, (m, 3)),
        "v": rng.normal(0, 0.5, (m, 3)),
        "dv_prev": rng.normal(0, 0.2, (m, 3)),
        "phi_normal": np.concatenate([rng.normal(0, 1, (m, 3)), rng.uniform(-0.01, 0.02, (m, 1))], axis=1),
    }
    target = rng.normal(0, 0.3, (m, 3))
    lay = learn.param_layout()
    p = rng.normal(0, 0.15, lay["_total"][0])
    return rs, nbr, rest2, frame, target, p


def test_scene_graph_matches_brute_force():
    sc = scenes.c3_tube(nx=5, ny=4, spacing=0.01)
    rs, nbr, _ = learn.scene_graph(sc)
    n = sc.n
    ref = [set() for _ in range(n)]
    off = 0
    for g in sc.garments:
        for t in np.asarray(g.mesh.triangles):
            for a in range(3):
                i, j = int(t[a]) + off, int(t[(a + 1) % 3]) + off
                ref[i].add(j)
                ref[j].add(i)
        off += g.mesh.n
    for i in range(n):
        assert list(nbr[rs[i]:rs[i + 1]]) == sorted(ref[i])


def test_reference_backward_matches_finite_differences():
    rs, nbr, rest2, frame, target, p = _small_case()
    _, loss, G = mpnn_ref.forward_backward(p, rs, nbr, rest2, frame, target, bf=False)
    g = mpnn_ref.flatten(G)
    lay = learn.param_layout()
    rng = np.random.default_rng(7)
    checked = 0
    for name, (o, s) in lay.items():
        if name.startswith("_"):
            continue
        size = int(np.prod(s))
        idx = o + rng.choice(size, size=min(4, size), replace=False)
        for k in idx:
            if name == "We" and (k - o) // learn.HIDDEN >= 10:
                continue  # unused feature rows: zero gradient, checked below
            if name in ("Wd", "bd") and (k - o) % learn.OUT >= learn.OUT_USED:
                continue
            eps = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[k] += eps
            pm[k] -= eps
            lp = mpnn_ref.forward_backward(pp, rs, nbr, rest2, frame, target, bf=False)[1]
            lm = mpnn_ref.forward_backward(pm, rs, nbr, rest2, frame, target, bf=False)[1]
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (name, k - o, fd, g[k])
            checked += 1
    assert checked > 60
    # unused decoder columns and feature rows get no gradient
    o, s = lay["Wd"]
    assert not np.any(g[o:o + int(np.prod(s))].reshape(s)[:, learn.OUT_USED:])
    o, s = lay["We"]
    assert not np.any(g[o:o + int(np.prod(s))].reshape(s)[10:, :])


def test_bf16_rounding_is_round_to_nearest_even():
    # BF16 keeps 7 mantissa bits: 1 + 2^-7 is representable, 1 + 2^-8 is a tie between 1 and 1 + 2^-7.
    assert mpnn_ref.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0              # tie, 1.0 has even mantissa
    assert mpnn_ref.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6  # tie up to even
    assert mpnn_ref.bf16(np.array([1.0 + 2 ** -7]))[0] == 1.0 + 2 ** -7      # representable
    assert mpnn_ref.bf16(np.array([-2.5]))[0] == -2.5
