"""Pins of the oracle's adaptive light sampling: the weighted
AccurateSoftShadow average (AccurateSoftShadow.frag:79-87) and one quad-tree
node's test (QuadTree/Reprojection.frag, QuadTree/Evaluation.frag)."""
import numpy as np
import pytest

SI = 0.25


def _identity_trans(S, w=1.0):
    t = np.zeros((S, 4), np.float32)
    t[:, 3] = w
    return t


def _scene(n=64, seed=0):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0.05, 0.95, size=(n, 3)).astype(np.float32)
    pos = np.ascontiguousarray(pts.T.reshape(3, 1, n))
    m = np.eye(4, dtype=np.float32).reshape(-1).copy()
    m[12:16] = 0  # columns 0-2 only: the translation comes from trans
    return pts, pos, m


def test_adaptive_level_zero_equals_monte_carlo(oracle_lib):
    """Every sample at level 0 weighs 1/4: sum(v / 4) / 1 is bit-identical to
    sum(v) / 4 (a power-of-two scale), so the adaptive pass equals the
    Monte-Carlo pass on the same samples."""
    rng = np.random.default_rng(1)
    maps = (rng.random((4, 32, 32)) > 0.5).astype(np.float32)
    _, pos, m = _scene()
    trans = _identity_trans(4)
    trans[:, :2] = rng.uniform(-0.05, 0.05, size=(4, 2))
    a = oracle_lib.soft_shadow_adaptive(pos, maps, m, trans, si=SI)
    b = oracle_lib.soft_shadow(pos, maps, m, trans, si=SI)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()
    a = oracle_lib.soft_shadow_adaptive(pos, maps, m, trans, si=SI, pcf=True)
    b = oracle_lib.soft_shadow(pos, maps, m, trans, si=SI, pcf=True)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()


def test_adaptive_weighted_mean_closed_form(oracle_lib):
    """A lit sample at level 0 (weight 1/4) and a shadowed one at level 1
    (weight 1/9): (1/4 + si/9) / (1/4 + 1/9)."""
    maps = np.stack([np.ones((16, 16), np.float32), np.zeros((16, 16), np.float32)])
    _, pos, m = _scene(16)
    pos[2] = 0.5
    trans = _identity_trans(2)
    trans[1, 3] = np.float32(1.0) + np.float32(10000.0)
    got = oracle_lib.soft_shadow_adaptive(pos, maps, m, trans, si=SI)[0]
    np.testing.assert_allclose(got, (0.25 + SI / 9.0) / (0.25 + 1.0 / 9.0), rtol=1e-6)


def test_adaptive_level_clamp_decodes_with_level_four(oracle_lib):
    """Levels above 4 weigh as level 4, and the shader subtracts 4 * 10000
    from w, so a level-6 sample projects with w + 20000: one such sample
    equals the Monte-Carlo pass with that w."""
    rng = np.random.default_rng(2)
    maps = (rng.random((1, 32, 32)) > 0.5).astype(np.float32)
    _, pos, m = _scene()
    enc = _identity_trans(1, 1.0)
    enc[0, 3] = np.float32(1.0) + np.float32(60000.0)
    dec = _identity_trans(1)
    dec[0, 3] = np.float32(enc[0, 3] - np.float32(40000.0))
    a = oracle_lib.soft_shadow_adaptive(pos, maps, m, enc, si=SI)
    b = oracle_lib.soft_shadow(pos, maps, m, dec, si=SI)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()
    assert dec[0, 3] != 1.0


def _node_scene(w=8, h=6):
    """A G-buffer covering every pixel of a w x h view, depth 0.5."""
    ii, jj = np.meshgrid(np.arange(w), np.arange(h))
    pos = np.zeros((3, h, w), np.float32)
    pos[0] = (ii + 0.5) / w
    pos[1] = (jj + 0.5) / h
    pos[2] = 0.5
    return pos


def _node_mvps():
    m = np.eye(4, dtype=np.float32).reshape(-1)
    return np.stack([m] * 4)


@pytest.mark.parametrize("lit,want_sum", [(4, 4.0), (0, 0.0), (2, 2.0), (3, 3.0)])
def test_node_test_reprojection_counts_lit_corners(oracle_lib, lit, want_sum):
    pos = _node_scene()
    maps = np.stack([np.ones((16, 16), np.float32)] * lit + [np.zeros((16, 16), np.float32)] * (4 - lit))
    img, count = oracle_lib.qt_node_test(pos, maps, [0, 1, 2, 3], _node_mvps(), 1)
    np.testing.assert_array_equal(img, np.full(img.shape, want_sum, np.float32))
    # 0 or 4 lit corners: no penumbra; 1-3: every pixel whose 3x3 neighbourhood
    # stays inside the view (border texels read 0) passes
    h, w = img.shape
    assert count == (0 if lit in (0, 4) else (w - 2) * (h - 2))


def test_node_test_quarter_grid_and_background(oracle_lib):
    """div 2: the evaluation grid is (w/2) x (h/2) and samples texel 2c + 1;
    a background pixel (vertex.x == 0) reads 0.63, fails itself and every
    cell whose 3x3 neighbourhood samples it."""
    maps = np.stack([np.ones((16, 16), np.float32)] * 2 + [np.zeros((16, 16), np.float32)] * 2)
    pos = _node_scene(8, 6)
    img, count = oracle_lib.qt_node_test(pos, maps, [0, 1, 2, 3], _node_mvps(), 2)
    assert count == (4 - 2) * (3 - 2)  # cells (1, 1) and (2, 1)
    pos[0, 3, 2] = 0.0  # texel (2, 3): no cell samples an even column
    img, count = oracle_lib.qt_node_test(pos, maps, [0, 1, 2, 3], _node_mvps(), 2)
    assert img[3, 2] == np.float32(0.63) and count == 2
    # at full resolution the 3x3 block around it loses its 9 interior pixels
    img, count = oracle_lib.qt_node_test(pos, maps, [0, 1, 2, 3], _node_mvps(), 1)
    assert count == (8 - 2) * (6 - 2) - 9
    pos = _node_scene(8, 6)
    pos[0, 3, 3] = 0.0  # texel (3, 3), sampled by cell (1, 1), a neighbour of cell (2, 1)
    img, count = oracle_lib.qt_node_test(pos, maps, [0, 1, 2, 3], _node_mvps(), 2)
    assert count == 0
