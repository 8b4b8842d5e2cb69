"""Pins of the oracle's revectorization-based adaptive sampling
(ShadowParams.revectorizationBasedAdaptiveSampling):
HardShadow/Discontinuity.frag, QuadTree/RevectorizationBasedReprojection.frag
and SoftShadow/RevectorizationBasedAccurateSoftShadow.frag.

Synthetic G-buffers whose points project with the identity matrix, so a point
(x, y, z) reads texel (x W, y H) at depth z; normals face a light straight
above unless a test turns them away."""
import numpy as np
import pytest

SI = 0.25
N = 64
EYE4 = np.eye(4, dtype=np.float32).reshape(-1)


def _gbuf(pts, normal=(0.0, 0.0, 1.0)):
    pts = np.asarray(pts, np.float32).reshape(-1, 3)
    pos = np.ascontiguousarray(pts.T.reshape(3, 1, -1))
    nrm = np.zeros_like(pos)
    nrm[0], nrm[1], nrm[2] = normal
    return pos, nrm


def _params(oracle_lib, n=N, si=SI, mvp=EYE4):
    return oracle_lib.make_params(mvp, n, n, "naive", shadow_intensity=si, max_search=16, depth_threshold=2.5e-6,
                                  pre_evaluation=True, light_position=(0.5, 0.5, 100.0))


def _left_half(n=N, blocker=0.5):
    d = np.ones((n, n), np.float32)
    d[:, : n // 2] = blocker
    return d


def _texel(c, u=0.5, n=N):
    return (c + u) / n


# --------------------------------------------------------------------------
# Discontinuity.frag
# --------------------------------------------------------------------------
def test_discontinuity_codes_on_a_half_plane(oracle_lib):
    """Occluder left of x = 0.5 (texels 0..31), receiver z = 0.9: texel 31 is
    in umbra with a lit right neighbour (r = 0.25, discType 1), texel 32 is lit
    with a shadowed left neighbour (r = 0.5, discType 0), interior texels have
    no discontinuity but alpha 1 (and discType 1 in umbra), a background pixel
    keeps the clear colour and a pixel the pre-evaluation turns away stores 0."""
    d = _left_half()
    y = _texel(20)
    pts = [(_texel(31), y, 0.9), (_texel(32), y, 0.9), (_texel(10), y, 0.9), (_texel(50), y, 0.9),
           (0.0, 0.0, 0.0)]
    pos, nrm = _gbuf(pts)
    got = oracle_lib.rev_discontinuity(pos, nrm, d, _params(oracle_lib))[0]
    assert got.tolist() == [1 | 16 | 32, 2 | 32, 16 | 32, 32, 0xFF]
    pos, nrm = _gbuf(pts[:1], normal=(0.0, 0.0, -1.0))
    assert oracle_lib.rev_discontinuity(pos, nrm, d, _params(oracle_lib))[0].tolist() == [0]


def test_discontinuity_needs_positive_w(oracle_lib):
    """w <= 0 (a point behind the light) stores 0 whatever the map holds."""
    d = _left_half()
    m = np.eye(4, dtype=np.float32)
    m[3, 3] = -1.0  # column-major: w = -1
    pos, nrm = _gbuf([(_texel(31), _texel(20), 0.9)])
    got = oracle_lib.rev_discontinuity(pos, nrm, d, _params(oracle_lib, mvp=m.reshape(-1)))
    assert got.tolist() == [[0]]


# --------------------------------------------------------------------------
# RevectorizationBasedReprojection.frag + Evaluation.frag
# --------------------------------------------------------------------------
def _maps(lit):
    """Four uniform maps: 1.0 (lit at z = 0.9) or 0.5 (umbra)."""
    return np.stack([np.full((N, N), 1.0 if v else 0.5, np.float32) for v in lit])


def _field(h=8, w=8, z=0.9):
    ys, xs = np.mgrid[0:h, 0:w]
    pts = np.stack([(xs + 0.5) / w, (ys + 0.5) / h, np.full(xs.shape, z)], -1).astype(np.float32)
    pos = np.ascontiguousarray(np.moveaxis(pts, -1, 0))
    return pos


MVPS = np.tile(EYE4, 4)


@pytest.mark.parametrize("lit,want_sum,want_vis", [((1, 1, 1, 1), 4.0, 1.0), ((0, 0, 0, 0), 0.0, 0.0),
                                                    ((1, 0, 1, 0), 2.0, 0.5)])
def test_node_test_level0_classification(oracle_lib, lit, want_sum, want_vis):
    """Level 0 starts the classification: all four corners lit is 1, none lit
    0, anything else 0.5; the image holds the lit-corner count, and only the
    mixed case passes Evaluation.frag (interior pixels whose 3 x 3
    neighbourhood agrees: 6 x 6 of the 8 x 8 grid; the border reads 0)."""
    pos = _field()
    disc = np.zeros((50, 8, 8), np.uint8)
    vis = np.full((8, 8), 7.0, np.float32)
    img, count = oracle_lib.rev_node_test(pos, _maps(lit), [0, 1, 2, 3], MVPS, disc, 0, vis)
    assert (img == want_sum).all() and (vis == want_vis).all()
    assert count == (36 if want_sum == 2.0 else 0)


def test_node_test_two_discontinuities_force_sum_5(oracle_lib):
    """Two corner layers flagging a discontinuity at a pixel force the count
    to 5, which Evaluation.frag never passes; one flag does not."""
    pos = _field()
    disc = np.zeros((50, 8, 8), np.uint8)
    disc[0, 2, 3] = 1 | 32          # r = 0.25
    disc[1, 2, 3] = (2 << 2) | 32   # g = 0.5
    disc[2, 5, 5] = 3 | 16 | 32     # one layer only
    disc[3, 6, 6] = 32 | 16         # alpha and discType only: no discontinuity
    disc[1, 6, 6] = 32
    vis = np.zeros((8, 8), np.float32)
    img, _ = oracle_lib.rev_node_test(pos, _maps((1, 0, 1, 0)), [0, 1, 2, 3], MVPS, disc, 0, vis)
    assert img[2, 3] == 5.0 and vis[2, 3] == 0.5
    assert img[5, 5] == 2.0 and img[6, 6] == 2.0


def test_node_test_layers_clamp_to_the_last(oracle_lib):
    """Slots past the 50-layer array read layer 49 (texture2DArray clamps the
    layer): flags planted only in layer 49 count for corners 60..63."""
    pos = _field()
    maps = np.concatenate([np.ones((60, N, N), np.float32), _maps((1, 1, 0, 0))])
    disc = np.zeros((50, 8, 8), np.uint8)
    disc[49, 4, 4] = 1 | 32
    vis = np.zeros((8, 8), np.float32)
    img, _ = oracle_lib.rev_node_test(pos, maps, [60, 61, 62, 63], MVPS, disc, 0, vis)
    assert img[4, 4] == 5.0 and img[3, 3] == 2.0


@pytest.mark.parametrize("prev,lit,want", [(0.0, (0, 0, 0, 0), 0.0), (0.0, (1, 0, 0, 0), 0.5),
                                           (1.0, (1, 1, 1, 1), 1.0), (1.0, (1, 1, 1, 0), 0.5),
                                           (0.5, (1, 1, 1, 1), 0.5), (0.5, (0, 0, 0, 0), 0.5)])
def test_node_test_deeper_levels_only_demote(oracle_lib, prev, lit, want):
    """Below level 0 the classification only moves to 0.5: umbra stays umbra
    while all corners stay dark, lit while all stay lit; 0.5 stays 0.5."""
    pos = _field()
    disc = np.zeros((50, 8, 8), np.uint8)
    vis = np.full((8, 8), prev, np.float32)
    oracle_lib.rev_node_test(pos, _maps(lit), [0, 1, 2, 3], MVPS, disc, 2, vis, div=2)
    assert (vis == want).all()


def test_node_test_background_and_div(oracle_lib):
    """Background pixels take the clear colour (0.63) in both targets; with
    div 2 Evaluation.frag runs on the 4 x 4 grid, whose interior 2 x 2 passes."""
    pos = _field()
    pos[:, 0, 0] = 0.0
    disc = np.zeros((50, 8, 8), np.uint8)
    vis = np.zeros((8, 8), np.float32)
    img, count = oracle_lib.rev_node_test(pos, _maps((1, 0, 0, 0)), [0, 1, 2, 3], MVPS, disc, 0, vis, div=2)
    assert img[0, 0] == np.float32(0.63) and vis[0, 0] == np.float32(0.63)
    assert count == 4


# --------------------------------------------------------------------------
# RevectorizationBasedAccurateSoftShadow.frag
# --------------------------------------------------------------------------
def _trans(levels):
    t = np.zeros((len(levels), 4), np.float32)
    t[:, 3] = 1.0 + np.asarray(levels, np.float32) * 10000.0
    return t


def _rev(oracle_lib, pts, maps, levels, disc, vis, rpcf=False, si=SI, normal=(0.0, 0.0, 1.0)):
    pos, nrm = _gbuf(pts, normal)
    p = _params(oracle_lib, si=si)
    return oracle_lib.rev_soft(pos, nrm, maps, EYE4, _trans(levels), disc, np.asarray(vis, np.float32).reshape(1, -1),
                               p, rpcf=rpcf)[0]


def test_final_pass_takes_the_classification(oracle_lib):
    """Umbra (0) becomes si, lit (1) stays 1, a pixel the pre-evaluation
    turns away is si whatever its classification; only 0.5 runs the
    samples."""
    maps = np.ones((2, N, N), np.float32)
    disc = np.zeros((50, 1, 3), np.uint8)
    pts = [(0.5, 0.5, 0.9)] * 3
    got = _rev(oracle_lib, pts, maps, [0, 0], disc, [0.0, 1.0, 0.5])
    assert got.tolist() == [SI, 1.0, 1.0]
    got = _rev(oracle_lib, pts, maps, [0, 0], disc, [0.0, 1.0, 0.5], normal=(0.0, 0.0, -1.0))
    assert got.tolist() == [SI, SI, SI]


def test_final_pass_weights_by_level(oracle_lib):
    """Penumbra pixel, one lit level-0 sample and one shadowed level-1 sample,
    no discontinuity: (1/4 + si/9) / (1/4 + 1/9), as the plain adaptive pass."""
    maps = np.stack([np.ones((N, N), np.float32), np.full((N, N), 0.5, np.float32)])
    disc = np.zeros((50, 1, 1), np.uint8)
    got = _rev(oracle_lib, [(0.5, 0.5, 0.9)], maps, [0, 1], disc, [0.5])
    f = np.float32
    want = (f(1.0) * f(0.25) + f(SI) * f(1.0 / 9.0)) / (f(0.25) + f(1.0 / 9.0))
    assert abs(got[0] - want) <= 1e-7


@pytest.mark.parametrize("rpcf", [False, True])
def test_final_pass_straight_edge_keeps_hard_values(oracle_lib, rpcf):
    """Every sample sees the same half-plane occluder.  Taps on the straight
    edge are revectorized to their hard values, so the binary pass gives the
    hard test (texel 31: si, texel 32: 1) and RPCF gives the 3 x 3 average
    ((3 + 6 si)/9 on texel 31, (6 + 3 si)/9 on texel 32)."""
    d = _left_half()
    maps = np.stack([d, d, d])
    pts = [(_texel(31, 0.3), _texel(20), 0.9), (_texel(32, 0.7), _texel(20), 0.9)]
    pos, nrm = _gbuf(pts)
    disc1 = oracle_lib.rev_discontinuity(pos, nrm, d, _params(oracle_lib))
    disc = np.zeros((50, 1, 2), np.uint8)
    disc[:3] = disc1
    got = _rev(oracle_lib, pts, maps, [0, 1, 1], disc, [0.5, 0.5], rpcf=rpcf)
    want = [(3 + 6 * SI) / 9, (6 + 3 * SI) / 9] if rpcf else [SI, 1.0]
    np.testing.assert_allclose(got, want, rtol=1e-6)


def test_final_pass_one_texel_features(oracle_lib):
    """The accurate shader's early exits: a lit texel ringed by umbra
    (entering discontinuity in all four directions) fills to 0, so it is
    shadowed (si); an umbra texel ringed by light (exiting in all four
    directions) is left lit (1)."""
    hole = np.full((N, N), 0.5, np.float32)
    hole[20, 20] = 1.0
    speck = np.ones((N, N), np.float32)
    speck[20, 20] = 0.5
    for dmap, want in ((hole, SI), (speck, 1.0)):
        pts = [(_texel(20), _texel(20), 0.9)]
        pos, nrm = _gbuf(pts)
        code = oracle_lib.rev_discontinuity(pos, nrm, dmap, _params(oracle_lib))
        assert code[0, 0] & 15 == 15  # r = g = 0.75
        disc = np.zeros((50, 1, 1), np.uint8)
        disc[0] = code
        got = _rev(oracle_lib, pts, dmap[None], [0], disc, [0.5])
        assert got[0] == np.float32(want)


def test_final_pass_scan_reaches_the_map_border(oracle_lib):
    """A one-texel-wide umbra column of 8 texels (rows 28..35) ending inside
    the map: its middle texel has light on both sides (exiting, r = 0.75,
    g = 0).  The accurate shader scans rows +1, -1, +2, ... beside the column
    for umbra, finds it only where the scan leaves the map (border depth 0),
    and the column stays in umbra (si).  RBSSM.frag's single probe row finds
    none and lights the texel's right part (fill 1 - ux)."""
    d = np.ones((N, N), np.float32)
    d[28:36, 20] = 0.5
    pts = [(_texel(20, 0.25), _texel(31), 0.9)]
    pos, nrm = _gbuf(pts)
    code = oracle_lib.rev_discontinuity(pos, nrm, d, _params(oracle_lib))
    assert code[0, 0] == 3 | 16 | 32
    disc = np.zeros((50, 1, 1), np.uint8)
    disc[0] = code
    got = _rev(oracle_lib, pts, d[None], [0], disc, [0.5])
    assert got[0] == np.float32(SI)
    # the same tap through RBSSM's smoothONDS (one probe row)
    t = oracle_lib.rbssm_tap(d, _params(oracle_lib), pts[0][0], pts[0][1], 0.9)
    assert t["ny"] <= -2.0 and t["fill"] == np.float32(0.75)


def test_final_pass_shadow_intensity_half_quirk(oracle_lib):
    """With shadowIntensity 0.5 a pixel the pre-evaluation turns away carries
    0.5, which the shader takes for a penumbra pixel and samples."""
    maps = np.ones((1, N, N), np.float32)
    disc = np.zeros((50, 1, 1), np.uint8)
    got = _rev(oracle_lib, [(0.5, 0.5, 0.9)], maps, [0], disc, [0.0], si=0.5, normal=(0.0, 0.0, -1.0))
    assert got[0] == 1.0
