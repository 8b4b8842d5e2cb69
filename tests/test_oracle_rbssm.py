"""Pins of the oracle's RBSSM (SoftShadowMapping/Shaders/SoftShadow/RBSSM.frag).

What the geometry fixes:
  * a long straight shadow edge is left as it is: every tap on it keeps its
    hard value, so the pass equals PCSS's PCF over the same taps (the PCSS
    oracle is pinned by closed forms of its own);
  * the edge walks (computeDiscontinuityLength) measure, in texels, the
    distance to the first texel past the end of the edge: positive where the
    edge ends at a convex corner, negative where it stops at a step or on the
    lit side, -(maxSearch + 1) when no end is found;
  * normalizeDiscontinuitySpace places the tap's sub-texel point on the edge:
    with both ends found it is min(distance to an end) / edge length;
  * umbra, no-penumbra (zNear 0, blockers shallower than 0.99), a kernel step
    of one texture width or more, and background pixels have closed values.
The fill smoothONDS gives at edge ends and corners has no closed form; it is
transcribed branch by branch and the GPU kernel is checked against it."""
import numpy as np
import pytest

SI = 0.25


def _params(oracle_lib, depth, si=SI, max_search=16, thr=2.5e-6):
    h, w = depth.shape
    return oracle_lib.make_params(np.eye(4, dtype=np.float32).reshape(-1), w, h, "naive", shadow_intensity=si,
                                  max_search=max_search, depth_threshold=thr)


def _pass(oracle_lib, fn, depth, pts, zn=1, bss=7, ks=7, radius=8, si=SI):
    pts = np.asarray(pts, np.float32)
    pos = np.ascontiguousarray(pts.T.reshape(3, 1, -1))
    p = _params(oracle_lib, depth, si)
    return fn(pos, None, depth, p, blocker_search_size=bss, kernel_size=ks, light_source_radius=radius,
              z_near=zn)[0]


def _edge_map(orientation, n=1024, edge=512, blocker=0.995):
    d = np.ones((n, n), np.float32)
    if orientation == "left":
        d[:, :edge] = blocker
    elif orientation == "right":
        d[:, edge:] = blocker
    elif orientation == "below":
        d[:edge, :] = blocker
    else:
        d[edge:, :] = blocker
    return d


def test_rbssm_half_plane_closed_form(oracle_lib):
    """Half-plane occluder at 0.995 left of x = 0.5, receiver 0.999, radius 8,
    zNear 1, kernelSize 7: the penumbra is ~0.032, so of the 7 x 7 taps the
    three columns w < 0 are shadowed and the tap column w = 0 sits on texel
    512, the lit side of the edge, whose revectorized fill stays 1:
    (28 + 21 si) / 49.  Deep inside the occluder every tap is shadowed (si);
    far on the lit side no blocker is found (1)."""
    d = _edge_map("left")
    got = _pass(oracle_lib, oracle_lib.rbssm, d, [(0.5, 0.5, 0.999), (0.2, 0.5, 0.999), (0.8, 0.5, 0.999)])
    np.testing.assert_allclose(got, [(28 + 21 * SI) / 49, SI, 1.0], rtol=1e-6)


@pytest.mark.parametrize("orientation", ["left", "right", "below", "above"])
def test_rbssm_straight_edge_equals_pcf(oracle_lib, orientation):
    """Points within a penumbra of a straight edge, so that many taps land on
    the two texel rows beside it: bit-identical to PCSS's PCF."""
    d = _edge_map(orientation)
    rng = np.random.default_rng({"left": 1, "right": 2, "below": 3, "above": 4}[orientation])
    m = 300
    across = rng.uniform(0.5 - 0.03, 0.5 + 0.03, m)
    along = rng.uniform(0.3, 0.7, m)
    if orientation in ("left", "right"):
        pts = np.column_stack([across, along, np.full(m, 0.999)])
    else:
        pts = np.column_stack([along, across, np.full(m, 0.999)])
    pts = pts.astype(np.float32)
    got = _pass(oracle_lib, oracle_lib.rbssm, d, pts)
    want = _pass(oracle_lib, oracle_lib.pcss, d, pts, ks=7)
    np.testing.assert_array_equal(got, want)
    assert 0 < np.count_nonzero((got > SI) & (got < 1)), "no point in the penumbra"


@pytest.mark.parametrize("orientation", ["left", "right", "below", "above"])
def test_rbssm_straight_edge_taps_keep_hard_values(oracle_lib, orientation):
    """The two texels beside a long straight edge, at any sub-texel position:
    the walks along the edge find no end (-(maxSearch + 1) both ways), the
    walks across it stop at once (-1), both normalize to -1, and the fill is
    the hard value: 0 on the shadowed side, 1 on the lit side."""
    n = 64
    d = _edge_map(orientation, n=n, edge=32)
    p = _params(oracle_lib, d)
    rng = np.random.default_rng(7)
    shadowed = {"left": 31, "right": 32, "below": 31, "above": 32}[orientation]
    lit = {"left": 32, "right": 31, "below": 32, "above": 31}[orientation]
    for texel, fill, b in ((shadowed, 0.0, 1.0), (lit, 1.0, 0.0)):
        for _ in range(8):
            ua, ub = rng.integers(0, 16, 2) / 16.0
            along = 30 + ub
            if orientation in ("left", "right"):
                x, y = (texel + ua) / n, along / n
            else:
                x, y = along / n, (texel + ua) / n
            t = oracle_lib.rbssm_tap(d, p, x, y, 0.999)
            assert t["b"] == b
            walks = (t["down"], t["up"], t["left"], t["right"]) if orientation in ("left", "right") else \
                (t["left"], t["right"], t["down"], t["up"])
            assert walks == (-17.0, -17.0, -1.0, -1.0), (texel, t)
            assert (t["nx"], t["ny"]) == (-1.0, -1.0)
            assert t["fill"] == fill, (texel, t)


def _rect_map(n=64):
    """Occluder at 0.995 over columns 20..29, rows 10..39 of a 64 x 64 map,
    with a taller part (rows 10..41) over columns 30..39: the top edge of the
    first part ends at a convex corner on its left (column 19 is lit) and at a
    step on its right."""
    d = np.ones((n, n), np.float32)
    d[10:40, 20:30] = 0.995
    d[10:42, 30:40] = 0.995
    return d


def test_rbssm_edge_walk_lengths(oracle_lib):
    """Top row (39) of the lower part, tap at column c: the walk to the left
    reaches lit column 19 after c - 19 texels and finds a convex end (+); the
    walk to the right runs into the step at column 30 after 30 - c texels
    (the texel above is no longer lit) and stops without an end (-).  The
    vertical walks of a horizontal edge return -1 at once.  One row up, on the
    lit side, the left walk stops where the texel below turns lit: -(c - 19)."""
    n = 64
    d = _rect_map(n)
    p = _params(oracle_lib, d)
    for c in (20, 23, 25, 29):
        t = oracle_lib.rbssm_tap(d, p, (c + 0.5) / n, (39 + 0.5) / n, 0.999)
        # column 20 is the corner texel: its left neighbour is lit too
        assert (t["r"], t["g"], t["b"]) == (0.5 if c == 20 else 0.0, 0.5, 1.0)
        assert t["left"] == c - 19 and t["right"] == -(30 - c), (c, t)
        if c != 20:
            assert (t["down"], t["up"]) == (-1.0, -1.0)
        u = oracle_lib.rbssm_tap(d, p, (c + 0.5) / n, (40 + 0.5) / n, 0.999)
        # column 29's right neighbour in row 40 is the taller part
        assert (u["r"], u["g"], u["b"]) == (0.25 if c == 29 else 0.0, 0.25, 0.0)
        assert u["left"] == -(c - 19), (c, u)


def test_rbssm_edge_walk_respects_max_search(oracle_lib):
    """The right part's top row (41) is 10 texels long with convex ends on
    both sides (columns 29 and 40 are lit at that row), so maxSearch 16 finds
    both ends; maxSearch 3 finds neither from the middle and returns
    -(3 + 1) both ways."""
    n = 64
    d = _rect_map(n)
    t = oracle_lib.rbssm_tap(d, _params(oracle_lib, d), (35 + 0.5) / n, (41 + 0.5) / n, 0.999)
    assert (t["left"], t["right"]) == (6.0, 5.0)
    t = oracle_lib.rbssm_tap(d, _params(oracle_lib, d, max_search=3), (35 + 0.5) / n, (41 + 0.5) / n, 0.999)
    assert (t["left"], t["right"]) == (-4.0, -4.0)


@pytest.mark.parametrize("c,ux", [(30, 0.25), (32, 0.5), (35, 0.5), (35, 0.75), (37, 0.125), (39, 0.9375)])
def test_rbssm_normalized_position_on_an_edge(oracle_lib, c, ux):
    """Both ends found on the 10-texel edge of row 41 (columns 30..39): the
    normalized coordinate, stored compressed below -2, decompresses to the
    sub-texel point's distance to the nearer end over the edge length."""
    n = 64
    d = _rect_map(n)
    t = oracle_lib.rbssm_tap(d, _params(oracle_lib, d), (c + ux) / n, (41 + 0.5) / n, 0.999)
    assert t["nx"] <= -2.0
    nd = 0.5 - ((t["nx"] + 2.0) * -1.0) / 2.0
    xs = c + ux
    want = min(xs - 30.0, 40.0 - xs) / 10.0
    assert abs(nd - want) < 1e-6, (t, want)


def test_rbssm_umbra_and_lit_closed_forms(oracle_lib):
    """Uniform occluder at 0.995 over the whole map: every tap is shadowed
    and none is on a discontinuity, so the result is si exactly; zNear 0
    (the reference app's value) makes the penumbra 0 and every pixel lit;
    blockers shallower than 0.99 give no penumbra (lit); a radius so large
    that the kernel step reaches one texture width returns 1."""
    n = 256
    d = np.full((n, n), 0.995, np.float32)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(0.1, 0.9, 100), rng.uniform(0.1, 0.9, 100),
                           np.full(100, 0.999)]).astype(np.float32)
    np.testing.assert_array_equal(_pass(oracle_lib, oracle_lib.rbssm, d, pts), np.full(100, SI, np.float32))
    np.testing.assert_array_equal(_pass(oracle_lib, oracle_lib.rbssm, d, pts, zn=0), np.ones(100, np.float32))
    np.testing.assert_array_equal(_pass(oracle_lib, oracle_lib.rbssm, np.full((n, n), 0.5, np.float32), pts),
                                  np.ones(100, np.float32))
    np.testing.assert_array_equal(_pass(oracle_lib, oracle_lib.rbssm, d, pts, radius=4096),
                                  np.ones(100, np.float32))


def test_rbssm_background_and_rejects(oracle_lib):
    """Background pixels (x = 0 in the vertex map) are discarded (0); a
    kernel or search size below 2 divides by a zero filter width and is
    rejected."""
    d = _edge_map("left", n=64, edge=32)
    pos = np.zeros((3, 1, 2), np.float32)
    pos[:, 0, 1] = (0.2, 0.5, 0.999)
    p = _params(oracle_lib, d)
    out = oracle_lib.rbssm(pos, None, d, p)
    assert out[0, 0] == 0.0 and out[0, 1] == SI
    with pytest.raises(RuntimeError):
        oracle_lib.rbssm(pos, None, d, p, kernel_size=1)
    with pytest.raises(RuntimeError):
        oracle_lib.rbssm(pos, None, d, p, blocker_search_size=1)


# --------------------------------------------------------------------------
# smoothONDS (RBSSM.frag:311-1117) where an edge ends, turns a corner or is
# one texel wide: fills traced by hand through the shader's branches
# --------------------------------------------------------------------------
def _fill(oracle_lib, d, p, x, y):
    """(tap, FP32 sub-texel coordinates) of an RBSSM tap at receiver 0.999."""
    n = d.shape[0]
    xf, yf = np.float32(x / n), np.float32(y / n)
    t = oracle_lib.rbssm_tap(d, p, xf, yf, 0.999)
    xs, ys = float(xf) * n, float(yf) * n
    return t, xs - np.floor(xs), ys - np.floor(ys)


def _cl(v):
    return min(max(v, 0.0), 1.0)


def test_smoothonds_revectorizes_an_edge_between_its_ends(oracle_lib):
    """The lower part's top edge runs along row 39 from a convex corner at
    column 20 (column 19 is lit) to a step at column 30 (the taller part).
    Umbra tap (c, 39), c = 21..29 (r = 0, g = .5, b = 1): the left walk
    reaches lit column 19 and finds the end, +(c - 19); the right walk loses
    the lit texel above at column 30, -(30 - c); both vertical walks return
    -1 at once (r = 0, g != 0).  normalizeDiscontinuitySpace gives
    len = 10 and nx = 1 - (c - 19)/10 + (1 - ux)/10; the exiting branch for a
    bottom discontinuity returns clamp(nx - (1 - uy)).  Lit tap (c, 40)
    (r = 0, g = .25, b = 0): walks -(c - 19) and +(30 - c) (the right walk
    enters umbra at the step), nx = 1 - (30 - c)/10 + ux/10, and the
    entering branch for a top discontinuity returns clamp(uy - (nx - 1))
    (column 29 of row 40 also borders the taller part on its right, a corner
    case of its own).
    Both are clamp(Y - 37 - X/10) in texel coordinates: the staircase end
    (X, Y) = (20, 39) .. (30, 40) is filled along one straight line."""
    n = 64
    d = _rect_map(n)
    p = _params(oracle_lib, d)
    rng = np.random.default_rng(31)
    for c in range(21, 30):
        for row in ((39, 40) if c < 29 else (39,)):
            for ux, uy in rng.uniform(0.02, 0.98, size=(6, 2)):
                t, fx, fy = _fill(oracle_lib, d, p, c + ux, row + uy)
                X, Y = c + fx, row + fy
                if row == 39:
                    assert (t["r"], t["g"], t["b"]) == (0.0, 0.5, 1.0)
                    assert (t["left"], t["right"]) == (c - 19, -(30 - c))
                else:
                    assert (t["r"], t["g"], t["b"]) == (0.0, 0.25, 0.0)
                    assert (t["left"], t["right"]) == (-(c - 19), 30 - c)
                assert (t["down"], t["up"]) == (-1.0, -1.0)
                assert abs(t["fill"] - _cl(Y - 37.0 - X / 10.0)) <= 2e-6, (c, row, ux, uy, t)


def test_smoothonds_convex_corner(oracle_lib):
    """Corner texel (20, 39) of the lower part: umbra with lit neighbours left
    and above, so r = g = .5, b = 1.  Walks: left and up return 1 at once
    (the neighbour in that direction is the discontinuity), right runs to the
    step (-10), down follows the lit column 19 for maxSearch steps (-17).
    normalize: nx = 1 - ux/10 (len 10), ny = 1 - (1 - uy)/16 (len 16).  The
    corner branch moves the tap two columns left (the shader adds the x step
    twice) and one row down before its dominant-axis tests; both find an
    exiting discontinuity, and the left-bottom case returns
    max(clamp(nx - (1 - uy)), clamp(ny - ux))."""
    n = 64
    d = _rect_map(n)
    p = _params(oracle_lib, d)
    rng = np.random.default_rng(32)
    for ux, uy in rng.uniform(0.02, 0.98, size=(40, 2)):
        t, fx, fy = _fill(oracle_lib, d, p, 20 + ux, 39 + uy)
        assert (t["r"], t["g"], t["b"]) == (0.5, 0.5, 1.0)
        assert (t["left"], t["right"], t["down"], t["up"]) == (1.0, -10.0, -17.0, 1.0)
        nx, ny = 1.0 - fx / 10.0, 1.0 - (1.0 - fy) / 16.0
        assert abs(t["nx"] - nx) <= 1e-6 and abs(t["ny"] - ny) <= 1e-6
        assert abs(t["fill"] - max(_cl(nx - (1.0 - fy)), _cl(ny - fx))) <= 2e-6, (ux, uy, t)


def _line_map(n=64, row=30, c0=10, c1=60):
    """A one-texel umbra line (0.995) along `row`, columns c0..c1, in a lit map."""
    d = np.ones((n, n), np.float32)
    d[row, c0:c1 + 1] = 0.995
    return d


def test_smoothonds_one_texel_line(oracle_lib):
    """Umbra tap on a one-texel line: lit above and below, so r = 0,
    g = .75, b = 1 (exiting top and bottom only).  The neighbours above and
    below have no discontinuity of their own along x, nor do those of the
    next column, so the branch (without the accurate shader's scan) falls to
    its last case: mix(1 - uy, clamp(nx - uy), step(-2, nx)).
    * Column 20 of a line over columns 10..60: the left walk finds the end at
      column 9, +11; the right walk runs maxSearch steps, -17; len 16 and
      nx = (6 - ux)/16, so the fill is clamp((6 - ux)/16 - uy).
    * Column 25 of a line over columns 10..40: both ends 16 texels away, both
      walks +16, nx compressed below -2, so the fill is 1 - uy."""
    rng = np.random.default_rng(33)
    d = _line_map()
    p = _params(oracle_lib, d)
    for ux, uy in rng.uniform(0.02, 0.98, size=(30, 2)):
        t, fx, fy = _fill(oracle_lib, d, p, 20 + ux, 30 + uy)
        assert (t["r"], t["g"], t["b"]) == (0.0, 0.75, 1.0)
        assert (t["left"], t["right"], t["down"], t["up"]) == (11.0, -17.0, -1.0, -1.0)
        assert abs(t["nx"] - (6.0 - fx) / 16.0) <= 1e-6
        assert abs(t["fill"] - _cl((6.0 - fx) / 16.0 - fy)) <= 2e-6, (ux, uy, t)
    d = _line_map(c1=40)
    p = _params(oracle_lib, d)
    for ux, uy in rng.uniform(0.02, 0.98, size=(30, 2)):
        t, fx, fy = _fill(oracle_lib, d, p, 25 + ux, 30 + uy)
        assert (t["left"], t["right"]) == (16.0, 16.0)
        assert t["nx"] <= -2.0
        assert abs(t["fill"] - (1.0 - fy)) <= 2e-6, (ux, uy, t)
