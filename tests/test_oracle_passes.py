"""Pins of the oracle's screen passes (a8-a14) against closed forms,
invariants and tiny brute-force cases.

Shadow coordinates are driven directly: with the identity as B*lightMVP, a
G-buffer position (x, y, z) projects to texture coordinate (x, y) and depth z,
so every expected value below is computed by hand from the map contents.
"""
import math

import numpy as np
import pytest

SI = 0.25
W = H = 16


def run_hard(oracle_lib, depth, pts, technique="naive", mvp=None, **kw):
    pts = np.asarray(pts, dtype=np.float32)
    n = len(pts)
    pos = np.zeros((3, 1, n), np.float32)
    pos[:, 0, :] = pts.T
    nrm = kw.pop("normals", None)
    if nrm is not None:
        nrm = np.asarray(nrm, np.float32).T.reshape(3, 1, n).copy()
    m = np.eye(4, dtype=np.float32).reshape(-1) if mvp is None else mvp
    h, w = depth.shape
    params = oracle_lib.make_params(m, w, h, technique, shadow_intensity=kw.pop("si", SI), **kw)
    return oracle_lib.hard_shadow(pos, nrm, depth, params)[0]


def test_fetch_nearest_and_border(oracle_lib):
    d = np.arange(W * H, dtype=np.float32).reshape(H, W) + 1.0
    for i, j in [(0, 0), (3, 7), (15, 15)]:
        for fx, fy in [(0.0, 0.0), (0.5, 0.5), (0.999, 0.001)]:
            assert oracle_lib.fetch(d, (i + fx) / W, (j + fy) / H) == d[j, i]
    # CLAMP_TO_BORDER with border depth 0: outside [0, 1) reads 0
    for s, t in [(-1e-6, 0.5), (1.0, 0.5), (0.5, 1.0), (0.5, -0.25), (float("nan"), 0.5)]:
        assert oracle_lib.fetch(d, s, t) == 0.0


def test_naive_closed_form(oracle_lib):
    rng = np.random.default_rng(0)
    d = rng.uniform(0.2, 0.8, size=(H, W)).astype(np.float32)
    pts = rng.uniform(0.01, 0.99, size=(200, 3)).astype(np.float32)
    got = run_hard(oracle_lib, d, pts)
    i = np.floor(pts[:, 0] * W).astype(int)
    j = np.floor(pts[:, 1] * H).astype(int)
    want = np.where(pts[:, 2] <= d[j, i], 1.0, SI)
    np.testing.assert_array_equal(got, want)


def test_naive_special_cases(oracle_lib):
    d = np.full((H, W), 0.5, np.float32)
    # x == 0 is background (Shadow.frag:244): the pass writes 0
    assert run_hard(oracle_lib, d, [(0.0, 0.5, 0.1)])[0] == 0.0
    # outside the map: border depth 0 -> shadowed
    assert run_hard(oracle_lib, d, [(1.25, 0.5, 0.1)])[0] == SI
    # w <= 0 (behind the light): lit without a test (Shadow.frag:251)
    m = np.eye(4, dtype=np.float32)
    m[3, 3] = -1.0  # column-major index 15: w = -1
    assert run_hard(oracle_lib, d, [(0.5, 0.5, 0.9)], mvp=m.reshape(-1))[0] == 1.0
    # depth equal to the stored value is lit (<=)
    assert run_hard(oracle_lib, d, [(0.5, 0.5, 0.5)])[0] == 1.0


def pcf_taps(k, offset, inclusive):
    """Tap offsets of the PCF loop in float32 (Shadow.frag:93-99)."""
    off = np.float32(offset)
    step = np.float32(np.float32(2.0) * off / np.float32(k))
    w, taps = -off, []
    while (w <= off) if inclusive else (w < off):
        taps.append(float(w))
        w = np.float32(w + step)
    return taps


def test_pcf_tap_tables():
    assert len(pcf_taps(3, 1, False)) == 3 and len(pcf_taps(3, 1, True)) == 4
    np.testing.assert_allclose(pcf_taps(3, 1, False), [-1, -1 / 3, 1 / 3], atol=1e-6)
    np.testing.assert_allclose(pcf_taps(3, 1, True), [-1, -1 / 3, 1 / 3, 1], atol=1e-6)
    assert len(pcf_taps(7, 1, False)) == 7


@pytest.mark.parametrize("fx", [0.1, 0.4, 0.6, 0.9])
@pytest.mark.parametrize("fy", [0.2, 0.7])
def test_pcf_counts_taps(oracle_lib, fx, fy):
    """3x3 PCF (k=3, offset 1, '<' loop) over a map whose only lit column is
    i0 and only lit row is j0: the result is the fraction of the 9 taps at
    offsets {-1, -1/3, +1/3} texels that land on a lit texel."""
    i0, j0 = 7, 9
    d = np.zeros((H, W), np.float32)
    d[:, i0] = 1.0
    d[j0, :] = 1.0
    x, y = (i0 + fx) / W, (j0 + fy) / H
    got = run_hard(oracle_lib, d, [(x, y, 0.5)], technique="pcf", kernel_order=3, penumbra_size=1)
    offs = [-1.0, -1.0 / 3.0, 1.0 / 3.0]
    lit = 0
    for a in offs:
        for b in offs:
            ti = math.floor((i0 + fx) + a)
            tj = math.floor((j0 + fy) + b)
            lit += (ti == i0) or (tj == j0)
    want = (lit * 1.0 + (9 - lit) * SI) / 9.0
    assert abs(got - want) < 1e-6


def test_pcf_uniform_maps(oracle_lib):
    pts = [(0.3, 0.6, 0.5), (0.71, 0.12, 0.5)]
    lit = run_hard(oracle_lib, np.ones((H, W), np.float32), pts, technique="pcf")
    dark = run_hard(oracle_lib, np.zeros((H, W), np.float32), pts, technique="pcf")
    np.testing.assert_allclose(lit, 1.0)
    np.testing.assert_allclose(dark, SI)


def test_pre_evaluation(oracle_lib):
    """Shadow.frag:222-238: N = normalize(normalMatrix * n) (pre-flipped for
    back faces), L = normalize(lightPosition - MV*v); dot(N, L) <= 0 gives si
    without a shadow-map test."""
    d = np.ones((H, W), np.float32)  # everything lit by the map
    pts = [(0.5, 0.5, 0.0)] * 3
    normals = [(0, 0, 1), (0, 0, -1), (1, 0, 0)]  # facing the light, away, perpendicular
    got = run_hard(oracle_lib, d, pts, technique="naive", pre_evaluation=True, normals=normals,
                   light_position=(0.5, 0.5, 10.0))
    np.testing.assert_array_equal(got, [1.0, SI, SI])


def test_soft_shadow_limits(oracle_lib):
    rng = np.random.default_rng(2)
    S = 4
    pos = np.zeros((3, 1, 8), np.float32)
    pos[0, 0] = rng.uniform(0.2, 0.8, 8)
    pos[1, 0] = rng.uniform(0.2, 0.8, 8)
    pos[2, 0] = 0.5
    m = np.eye(4, dtype=np.float32).reshape(-1)
    trans = np.tile(np.array([0, 0, 0, 1], np.float32), (S, 1))
    m_common = m.copy()
    m_common[12:16] = 0  # columns 0-2 only; column 3 comes from lightMVPTrans
    lit = oracle_lib.soft_shadow(pos, np.ones((S, H, W), np.float32), m_common, trans, si=SI)
    dark = oracle_lib.soft_shadow(pos, np.zeros((S, H, W), np.float32), m_common, trans, si=SI, pcf=True)
    np.testing.assert_allclose(lit, 1.0)
    np.testing.assert_allclose(dark, SI)


def test_soft_shadow_single_sample_equals_naive_and_mean(oracle_lib):
    """N = 1 reduces AccurateSoftShadow to the naive test; N > 1 is the mean
    of the per-sample tests (accFactor = 1)."""
    rng = np.random.default_rng(4)
    S = 5
    maps = rng.uniform(0.3, 0.7, size=(S, H, W)).astype(np.float32)
    pts = rng.uniform(0.05, 0.95, size=(64, 3)).astype(np.float32)
    pos = np.ascontiguousarray(pts.T.reshape(3, 1, -1))
    m = np.eye(4, dtype=np.float32).reshape(-1).copy()
    m[12:16] = 0
    offsets = rng.uniform(-0.05, 0.05, size=(S, 2)).astype(np.float32)
    trans = np.zeros((S, 4), np.float32)
    trans[:, :2] = offsets
    trans[:, 3] = 1.0
    per = []
    for s in range(S):
        one = oracle_lib.soft_shadow(pos, maps[s:s + 1], m, trans[s:s + 1], si=SI)[0]
        i = np.floor((pts[:, 0] + offsets[s, 0]) * W).astype(int)
        j = np.floor((pts[:, 1] + offsets[s, 1]) * H).astype(int)
        inside = (i >= 0) & (i < W) & (j >= 0) & (j < H)
        dd = np.where(inside, maps[s][np.clip(j, 0, H - 1), np.clip(i, 0, W - 1)], 0.0)
        np.testing.assert_array_equal(one, np.where(pts[:, 2] <= dd, 1.0, SI))
        per.append(one)
    allS = oracle_lib.soft_shadow(pos, maps, m, trans, si=SI)[0]
    np.testing.assert_allclose(allS, np.mean(per, axis=0), atol=2e-7)


# --------------------------------------------------------------------------
# RBSM (NonConservativeSMSR.frag / FilteredRBSM.frag)
# --------------------------------------------------------------------------
RW = 32


def staircase():
    ii, jj = np.meshgrid(np.arange(RW), np.arange(RW))
    d = np.ones((RW, RW), np.float32)
    d[(ii + jj) < 32] = 0.0  # texel (i, j) occluded when i + j <= 31
    return d


def rbsm_points(rng, n, near_line=True):
    pts = []
    while len(pts) < n:
        i = int(rng.integers(4, 28))
        j = 32 - i + int(rng.integers(-2, 2)) if near_line else int(rng.integers(2, 30))
        fx, fy = rng.random(2)
        pts.append((np.float32((i + fx) / RW), np.float32((j + fy) / RW), 0.5))
    return np.asarray(pts, np.float32)


def test_rbsm_revectorizes_45_degree_staircase(oracle_lib):
    """The 45-degree staircase is revectorized into the line through the step
    midpoints, x + y = 32.5 (texel units): binary RBSM classifies every point
    by that line, while the naive test follows the stairs."""
    rng = np.random.default_rng(0)
    pts = rbsm_points(rng, 3000)
    d = staircase()
    got = run_hard(oracle_lib, d, pts, technique="rbsm", si=0.0, max_search=16, depth_threshold=0.0)
    naive = run_hard(oracle_lib, d, pts, technique="naive", si=0.0)
    s = pts[:, 0] * RW + pts[:, 1] * RW
    far = np.abs(s - 32.5) > 0.05
    want = (s >= 32.5).astype(np.float32)
    np.testing.assert_array_equal(got[far], want[far])
    assert (naive[far] != want[far]).mean() > 0.02  # the plain test differs


def test_rbsm_axis_aligned_edge_is_unchanged(oracle_lib):
    # 64 x 64 map, samples at least maxSearch texels from the map border so
    # the edge walks never read the border (which would look like a corner)
    n = 64
    d = np.ones((n, n), np.float32)
    d[:, :32] = 0.0
    pts = np.array([((i + f) / n, (j + 0.5) / n, 0.5) for i in range(26, 38) for f in (0.05, 0.5, 0.95)
                    for j in (20, 33, 43)], np.float32)
    rb = run_hard(oracle_lib, d, pts, technique="rbsm", si=0.0, depth_threshold=0.0)
    nv = run_hard(oracle_lib, d, pts, technique="naive", si=0.0)
    np.testing.assert_array_equal(rb, nv)
    # RPCF of binary revectorized taps equals the plain average of the 4x4 taps
    rp = run_hard(oracle_lib, d, pts, technique="rpcf", si=0.0, kernel_order=3, penumbra_size=1)
    taps = [-1.0, -1.0 / 3.0, 1.0 / 3.0, 1.0]
    want = [np.mean([d[int(math.floor(y * n + b)), int(math.floor(x * n + a))] for a in taps for b in taps])
            for x, y, _ in pts]
    np.testing.assert_allclose(rp, want, atol=1e-6)
    # the fractional variant ramps monotonically across the edge
    fl = run_hard(oracle_lib, d, pts[::3], technique="rbsm_filtered", si=0.0, kernel_order=3, penumbra_size=1)
    assert (np.diff(fl) >= -1e-7).all() and fl[0] == 0.0 and fl[-1] == 1.0


def test_rbsm_no_discontinuity_equals_naive(oracle_lib):
    rng = np.random.default_rng(1)
    pts = rng.uniform(0.1, 0.9, size=(100, 3)).astype(np.float32)
    for val in (0.0, 1.0):
        d = np.full((RW, RW), val, np.float32)
        for tech in ("rbsm", "rpcf", "rbsm_filtered"):
            got = run_hard(oracle_lib, d, pts, technique=tech, kernel_order=3, penumbra_size=1)
            want = np.where(pts[:, 2] <= val, 1.0, SI)
            np.testing.assert_allclose(got, want, atol=1e-7)


def test_rbsm_binary_transpose_symmetry(oracle_lib):
    """With depthThreshold = 0 the four edge walks treat x and y alike, so the
    transposed map sampled at swapped coordinates gives the same result."""
    rng = np.random.default_rng(2)
    d = staircase()
    d[5:9, 20:25] = 0.0
    d[20:27, 3:6] = 0.0
    pts = np.array([((rng.integers(0, RW) + (rng.integers(0, 16) + 0.5) / 16) / RW,
                     (rng.integers(0, RW) + (rng.integers(0, 16) + 0.5) / 16) / RW, 0.5) for _ in range(2000)],
                   np.float32)
    a = run_hard(oracle_lib, d, pts, technique="rbsm", depth_threshold=0.0)
    b = run_hard(oracle_lib, np.ascontiguousarray(d.T), pts[:, [1, 0, 2]], technique="rbsm", depth_threshold=0.0)
    np.testing.assert_array_equal(a, b)


def test_rbsm_values_in_range(oracle_lib):
    rng = np.random.default_rng(3)
    d = (rng.random((RW, RW)) > 0.5).astype(np.float32)
    pts = rng.uniform(0.05, 0.95, size=(500, 3)).astype(np.float32)
    pts[:, 2] = 0.5
    for tech in ("rbsm", "rpcf", "rbsm_filtered"):
        got = run_hard(oracle_lib, d, pts, technique=tech, kernel_order=3, penumbra_size=1)
        assert got.min() >= SI - 1e-7 and got.max() <= 1.0 + 1e-7
        if tech == "rbsm":
            assert set(np.unique(got).tolist()) <= {SI, 1.0}


def test_loops_that_cannot_advance_are_rejected(oracle_lib):
    d = np.ones((4, 4), np.float32)
    with pytest.raises(RuntimeError):
        run_hard(oracle_lib, d, [(0.5, 0.5, 0.5)], technique="pcf", penumbra_size=0)


# --------------------------------------------------------------------------
# Conservative RBSM (ConservativeSMSR.frag)
# --------------------------------------------------------------------------
def test_conservative_rbsm_staircase_line(oracle_lib):
    """Hand-derived for the 45-degree staircase (umbra where i + j <= 31): a lit
    texel on the first lit diagonal (i + j = 32) has umbra left and below; its
    left and bottom walks hit umbra at once (+fx, +fy), its right and top walks
    stop after one step without a discontinuity (-(1-fx), -(1-fy)); so
    r = (fx, fy) and the texel is shadowed where fx + fy < 1.  The conservative
    edge is the line x + y = 33 in texel units, half a texel diagonal outside
    the non-conservative 32.5."""
    rng = np.random.default_rng(10)
    pts = rbsm_points(rng, 3000)
    d = staircase()
    got = run_hard(oracle_lib, d, pts, technique="rbsm_conservative", si=0.0, max_search=16, depth_threshold=0.0)
    s = pts[:, 0] * RW + pts[:, 1] * RW
    far = np.abs(s - 33.0) > 0.05
    np.testing.assert_array_equal(got[far], (s[far] >= 33.0).astype(np.float32))
    nonc = run_hard(oracle_lib, d, pts, technique="rbsm", si=0.0, max_search=16, depth_threshold=0.0)
    band = (s > 32.6) & (s < 32.95)
    assert band.sum() > 50 and (nonc[band] == 1.0).all() and (got[band] == 0.0).all()


def test_conservative_rbsm_only_grows_the_shadow(oracle_lib):
    """Conservative: a texel the plain test shadows stays shadowed, so the
    result never exceeds the naive one; values are si or 1."""
    rng = np.random.default_rng(11)
    d = (rng.random((RW, RW)) > 0.6).astype(np.float32)
    pts = rng.uniform(0.05, 0.95, size=(3000, 3)).astype(np.float32)
    pts[:, 2] = 0.5
    got = run_hard(oracle_lib, d, pts, technique="rbsm_conservative", depth_threshold=0.0)
    naive = run_hard(oracle_lib, d, pts, technique="naive")
    assert (got <= naive).all()
    assert set(np.unique(got).tolist()) <= {SI, 1.0}
    assert (got < naive).mean() > 0.01


def test_conservative_rbsm_axis_aligned_edge_and_uniform_maps(oracle_lib):
    """A straight edge: the walks along it never find an end, T = 0 makes that
    axis 0 and rx * ry = 0, so the result equals the naive test; RPCF equals
    the plain average of its taps.  Uniform maps equal the naive test."""
    n = 64
    d = np.ones((n, n), np.float32)
    d[:, :32] = 0.0
    pts = np.array([((i + f) / n, (j + 0.5) / n, 0.5) for i in range(26, 38) for f in (0.05, 0.5, 0.95)
                    for j in (20, 33, 43)], np.float32)
    rb = run_hard(oracle_lib, d, pts, technique="rbsm_conservative", si=0.0, depth_threshold=0.0)
    nv = run_hard(oracle_lib, d, pts, technique="naive", si=0.0)
    np.testing.assert_array_equal(rb, nv)
    rp = run_hard(oracle_lib, d, pts, technique="rpcf_conservative", si=0.0, kernel_order=3, penumbra_size=1)
    taps = [-1.0, -1.0 / 3.0, 1.0 / 3.0, 1.0]
    want = [np.mean([d[int(math.floor(y * n + b)), int(math.floor(x * n + a))] for a in taps for b in taps])
            for x, y, _ in pts]
    np.testing.assert_allclose(rp, want, atol=1e-6)
    rng = np.random.default_rng(12)
    upts = rng.uniform(0.1, 0.9, size=(100, 3)).astype(np.float32)
    for val in (0.0, 1.0):
        u = np.full((RW, RW), val, np.float32)
        for tech in ("rbsm_conservative", "rpcf_conservative"):
            got = run_hard(oracle_lib, u, upts, technique=tech, kernel_order=3, penumbra_size=1)
            np.testing.assert_allclose(got, np.where(upts[:, 2] <= val, 1.0, SI), atol=1e-7)


def test_conservative_rpcf_is_the_tap_mean_on_the_staircase(oracle_lib):
    """On the staircase every lit texel next to umbra has a left or bottom
    umbra neighbour and never umbra on both sides, so each RPCF tap takes the
    same branch as the SMSR test at that tap: RPCF is the mean of SMSR over the
    4 x 4 taps (coordinates s + w / W formed in FP32, as the loop forms them)."""
    rng = np.random.default_rng(13)
    pts = rbsm_points(rng, 300)
    d = staircase()
    rp = run_hard(oracle_lib, d, pts, technique="rpcf_conservative", max_search=16, depth_threshold=0.0,
                  kernel_order=3, penumbra_size=1)
    incr = np.float32(1.0) / np.float32(RW)
    step = np.float32(2.0) * np.float32(1.0) / np.float32(3.0)
    ws = []
    w = np.float32(-1.0)
    while w <= np.float32(1.0):
        ws.append(w)
        w = np.float32(w + step)
    assert len(ws) == 4
    taps = np.array([(np.float32(x + a * incr), np.float32(y + b * incr), z) for x, y, z in pts for a in ws
                     for b in ws], np.float32)
    sm = run_hard(oracle_lib, d, taps, technique="rbsm_conservative", max_search=16, depth_threshold=0.0)
    want = sm.reshape(len(pts), 16).sum(axis=1, dtype=np.float32) / np.float32(16.0)
    np.testing.assert_allclose(rp, want, atol=1e-6)


def test_conservative_rbsm_transpose_symmetry(oracle_lib):
    rng = np.random.default_rng(14)
    d = staircase()
    d[5:9, 20:25] = 0.0
    d[20:27, 3:6] = 0.0
    pts = np.array([((rng.integers(0, RW) + (rng.integers(0, 16) + 0.5) / 16) / RW,
                     (rng.integers(0, RW) + (rng.integers(0, 16) + 0.5) / 16) / RW, 0.5) for _ in range(2000)],
                   np.float32)
    a = run_hard(oracle_lib, d, pts, technique="rbsm_conservative", depth_threshold=0.0)
    b = run_hard(oracle_lib, np.ascontiguousarray(d.T), pts[:, [1, 0, 2]], technique="rbsm_conservative",
                 depth_threshold=0.0)
    np.testing.assert_array_equal(a, b)


# --------------------------------------------------------------------------
# RBSM taps traced by hand through the shader (or_rbsm_tap returns the walks,
# the relative positions and revectorizeShadow's value of one tap)
# --------------------------------------------------------------------------
def _tap_params(oracle_lib, d, thr=0.0, max_search=16):
    h, w = d.shape
    return oracle_lib.make_params(np.eye(4, dtype=np.float32).reshape(-1), w, h, "naive", shadow_intensity=0.0,
                                  max_search=max_search, depth_threshold=thr)


def _subs(rng, n):
    """Sub-texel positions away from 0 (where rx * ry == 0 takes its own branch)."""
    return rng.uniform(0.02, 0.98, size=(n, 2))


def test_filtered_rbsm_staircase_coverage_closed_form(oracle_lib):
    """45-degree staircase (umbra where i + j <= 31), receiver 0.5, tap in
    texel (i, j) at sub-texel (fx, fy), traced through NonConservativeSMSR's
    walks (FilteredRBSM.frag shares them):
    * first lit diagonal (i + j = 32): r = .5, g = .25, b = 0; the left and
      down walks step into umbra and find an end at once (+fx, +fy), the right
      and up walks lose the discontinuity after one step (-(1-fx), -(1-fy));
      T = 1 on both axes, so rel = (fx, fy);
    * last umbra diagonal (i + j = 31): r = .25, g = .5, b = 1; walks
      (-fx, 1-fx, -fy, 1-fy), so rel = (1-fx, 1-fy).
    FilteredRBSM.frag:266-272 (same-sign branch, clamp(1-s + (2s-1)(rx+ry)))
    then gives min(fx+fy, 1) on the lit diagonal and max(fx+fy-1, 0) on the
    umbra one: both are clamp(x + y - 32, 0, 1) in texel units, the coverage
    of a unit-wide linear ramp centred on x + y = 32.5.  The binary shader
    thresholds the same sum at 1/2: lit iff x + y >= 32.5.  Off the two
    diagonals there is no discontinuity and the tap keeps its hard value."""
    d = staircase()
    p = _tap_params(oracle_lib, d)
    rng = np.random.default_rng(21)
    for fx, fy in _subs(rng, 40):
        for i, j in ((14, 18), (20, 12), (9, 23)):       # i + j = 32, lit
            for di, want_b in ((0, 0.0), (-1, 1.0)):     # and the umbra texel left of it
                x = np.float32((i + di + fx) / RW)
                y = np.float32((j + fy) / RW)
                xs, ys = float(x) * RW, float(y) * RW     # the tap's texel coordinates in FP32
                for filtered in (True, False):
                    t = oracle_lib.rbsm_tap(d, p, x, y, 0.5, filtered=filtered)
                    ufx, ufy = xs - math.floor(xs), ys - math.floor(ys)
                    assert t["b"] == want_b
                    if want_b == 0.0:
                        assert (t["r"], t["g"]) == (0.5, 0.25)
                        np.testing.assert_allclose([t["left"], t["right"], t["down"], t["up"]],
                                                   [ufx, -(1 - ufx), ufy, -(1 - ufy)], atol=2e-6)
                        np.testing.assert_allclose([t["rx"], t["ry"]], [ufx, ufy], atol=2e-6)
                    else:
                        assert (t["r"], t["g"]) == (0.25, 0.5)
                        np.testing.assert_allclose([t["left"], t["right"], t["down"], t["up"]],
                                                   [-ufx, 1 - ufx, -ufy, 1 - ufy], atol=2e-6)
                        np.testing.assert_allclose([t["rx"], t["ry"]], [1 - ufx, 1 - ufy], atol=2e-6)
                    ramp = min(max(xs + ys - 32.0, 0.0), 1.0)
                    if filtered:
                        assert abs(t["value"] - ramp) <= 4e-6, (i + di, j, fx, fy, t)
                    elif abs(xs + ys - 32.5) > 1e-5:
                        assert t["value"] == (1.0 if xs + ys >= 32.5 else 0.0)
    # two texels away from the diagonals: no discontinuity, hard values
    for i, j, want in ((13, 17, 0.0), (15, 18, 1.0)):
        t = oracle_lib.rbsm_tap(d, p, (i + 0.4) / RW, (j + 0.6) / RW, 0.5, filtered=True)
        assert (t["r"], t["g"], t["value"]) == (0.0, 0.0, want)


def _notch():
    """Umbra rows 0..20 except a lit notch at (10..12, 20); lit above."""
    d = np.ones((RW, RW), np.float32)
    d[:21, :] = 0.0
    d[20, 10:13] = 1.0
    return d


def test_filtered_rbsm_notch_takes_the_both_ends_branch(oracle_lib):
    """Middle texel (11, 20) of a three-texel lit notch in an umbra edge, traced
    by hand: umbra only below it, so r = 0, g = .25, b = 0.  The left walk
    crosses lit (10, 20), whose lower neighbour is umbra, and stops in umbra
    texel 9 with an end found: +(1 + fx); the right walk likewise
    +(1 + (1 - fx)); the down walk finds umbra at once (+fy); the up walk has
    no discontinuity to follow (-(1 - fy)).  Both x walks positive: T = -2 and
    edgeLength = 3, so FilteredRBSM.frag:234-243 gives
    rel_x = -2 |max(-2(1+fx), -2(2-fx))| / 3 = -4 min(1+fx, 2-fx) / 3, rel_y = fy
    (T = 1).  The signs differ, so revectorizeShadow returns
    (1 - s) + (2s - 1) max(rx, ry) = fy for the lit tap (s = 1).  The binary
    shader scales |max(T dx, T dy)| by max(T, 2s-1) = 1 instead, so
    rel_x = +2 min(1+fx, 2-fx) / 3: both positive, and the tap stays lit."""
    d = _notch()
    p = _tap_params(oracle_lib, d)
    rng = np.random.default_rng(22)
    for fx, fy in _subs(rng, 40):
        x, y = np.float32((11 + fx) / RW), np.float32((20 + fy) / RW)
        xs, ys = float(x) * RW, float(y) * RW
        ufx, ufy = xs - 11.0, ys - 20.0
        ft = oracle_lib.rbsm_tap(d, p, x, y, 0.5, filtered=True)
        bt = oracle_lib.rbsm_tap(d, p, x, y, 0.5, filtered=False)
        for t in (ft, bt):
            assert (t["r"], t["g"], t["b"]) == (0.0, 0.25, 0.0)
            np.testing.assert_allclose([t["left"], t["right"], t["down"], t["up"]],
                                       [1 + ufx, 2 - ufx, ufy, -(1 - ufy)], atol=4e-6)
            assert abs(t["ry"] - ufy) <= 4e-6
        m = min(1 + ufx, 2 - ufx)
        assert abs(ft["rx"] + 4.0 * m / 3.0) <= 4e-6
        assert abs(bt["rx"] - 2.0 * m / 3.0) <= 4e-6
        assert abs(ft["value"] - ufy) <= 4e-6, (fx, fy, ft)
        assert bt["value"] == 1.0


def _step_map(umbra_right: bool):
    """Vertical step at column 16 with both sides within a depth threshold
    0.01 of the receiver 0.5: columns < 16 at 0.504 (lit), >= 16 at 0.495
    (umbra), or swapped."""
    d = np.full((RW, RW), 0.504, np.float32)
    if umbra_right:
        d[:, 16:] = 0.495
    else:
        d[:, :16] = 0.495
    return d


def test_rbsm_depth_threshold_carries_new_depth(oracle_lib):
    """Umbra tap in column 16 beside lit column 15, depthThreshold 0.01,
    traced through NonConservativeSMSR.frag:96-232: r = .5, g = 0, b = 1.
    The down walk steps to (16, 15), still umbra; getDisc(vec4) reads the
    left neighbour 0.504, within the threshold of 0.5, so it lowers the depth
    to 0.49, records it in newDepth, and the left test is lit: the edge goes
    on.  The walk then carries newDepth (:221-222): at (16, 14), 0.49 <= 0.495
    turns the texel lit, which is the end of the edge: +(1 + fy); the up walk
    likewise +(1 + (1 - fy)).  The left walk meets lit column 15 at once
    (+fx); the right walk has no discontinuity to follow (-(1 - fx)).  T = -2
    on y with edgeLength 3: binary rel_y = max(-2, -1) * 2 min(1+fy, 2-fy) / 3,
    filtered rel_y = -4 min(1+fy, 2-fy) / 3; rel_x = fx.  Binary: |rx| + |ry|
    > 1/2 shadows the tap (0); filtered: signs differ, so
    1 - max(rx, ry) = 1 - fx.
    Without the carry the walk keeps depth 0.5, never sees the texel turn lit
    and runs maxSearch steps down to the map border: -(16 + fy)."""
    d = _step_map(umbra_right=True)
    p = _tap_params(oracle_lib, d, thr=0.01)
    rng = np.random.default_rng(23)
    for fx, fy in _subs(rng, 30):
        x, y = np.float32((16 + fx) / RW), np.float32((16 + fy) / RW)
        ufx, ufy = float(x) * RW - 16.0, float(y) * RW - 16.0
        m = min(1 + ufy, 2 - ufy)
        for filtered in (False, True):
            t = oracle_lib.rbsm_tap(d, p, x, y, 0.5, filtered=filtered)
            assert (t["r"], t["g"], t["b"]) == (0.5, 0.0, 1.0)
            np.testing.assert_allclose([t["left"], t["right"], t["down"], t["up"]],
                                       [ufx, -(1 - ufx), 1 + ufy, 2 - ufy], atol=4e-6)
            assert abs(t["rx"] - ufx) <= 4e-6
            if filtered:
                assert abs(t["ry"] + 4.0 * m / 3.0) <= 4e-6
                assert abs(t["value"] - (1.0 - ufx)) <= 4e-6
            else:
                assert abs(t["ry"] + 2.0 * m / 3.0) <= 4e-6
                assert t["value"] == 0.0
    # with threshold 0 nothing is nudged: the vertical walks follow the
    # straight edge, the down walk for maxSearch = 16 texels (-(16 + .5)),
    # the up walk until the border row 32 (depth 0) ends the discontinuity
    # after 15 texels (-(15 + .5))
    t0 = oracle_lib.rbsm_tap(d, _tap_params(oracle_lib, d, thr=0.0), (16.5) / RW, (16.5) / RW, 0.5)
    assert (t0["down"], t0["up"]) == (-16.5, -15.5)


def test_rbsm_depth_threshold_on_the_lit_side(oracle_lib):
    """Lit tap in column 16 beside umbra column 15, threshold 0.01: every
    texel a walk visits lies within the threshold of 0.5, so
    computeDiscontinuityLength lowers the walk's depth to 0.49 at its first
    step (:196-199) and tests that texel against 0.49.  The left walk's
    first texel (0.495) then reads lit, not umbra, and the walk stops without
    an end (-fx) where it would have found one (+fx) with threshold 0; the
    other three walks stop at once too: (-fx, -(1-fx), -fy, -(1-fy)).  Both
    axes T = 0, so rel = (0, 0) and the tap keeps its lit value."""
    d = _step_map(umbra_right=False)
    rng = np.random.default_rng(24)
    for fx, fy in _subs(rng, 20):
        x, y = np.float32((16 + fx) / RW), np.float32((16 + fy) / RW)
        ufx, ufy = float(x) * RW - 16.0, float(y) * RW - 16.0
        t = oracle_lib.rbsm_tap(d, _tap_params(oracle_lib, d, thr=0.01), x, y, 0.5)
        assert (t["r"], t["g"], t["b"]) == (0.5, 0.0, 0.0)
        np.testing.assert_allclose([t["left"], t["right"], t["down"], t["up"]],
                                   [-ufx, -(1 - ufx), -ufy, -(1 - ufy)], atol=4e-6)
        assert (t["rx"], t["ry"], t["value"]) == (0.0, 0.0, 1.0)
        t0 = oracle_lib.rbsm_tap(d, _tap_params(oracle_lib, d, thr=0.0), x, y, 0.5)
        assert abs(t0["left"] - ufx) <= 4e-6


def _rpcf_taps(x, y, n):
    incr = np.float32(1.0) / np.float32(n)
    step = np.float32(2.0) * np.float32(1.0) / np.float32(3.0)
    ws = []
    w = np.float32(-1.0)
    while w <= np.float32(1.0):
        ws.append(w)
        w = np.float32(w + step)
    return [(np.float32(x + a * incr), np.float32(y + b * incr)) for a in ws for b in ws]


@pytest.mark.parametrize("case", ["staircase", "notch", "threshold"])
def test_filtered_rpcf_is_the_mean_of_its_taps(oracle_lib, case):
    """FilteredRBSM.frag:306-349: the pass is the mean over the 4 x 4 FP32 tap
    coordinates of mix(revectorized tap, 1, si) (or the mixed hard value
    where a tap has no discontinuity); with the tap values pinned above this
    fixes the pass on the same maps."""
    d = {"staircase": staircase, "notch": _notch, "threshold": lambda: _step_map(True)}[case]()
    thr = 0.01 if case == "threshold" else 0.0
    centre = {"staircase": (15, 17), "notch": (11, 20), "threshold": (16, 16)}[case]
    rng = np.random.default_rng(25)
    si = np.float32(0.25)
    p = _tap_params(oracle_lib, d, thr=thr)
    pts = []
    for fx, fy in _subs(rng, 25):
        pts.append((np.float32((centre[0] + fx) / RW), np.float32((centre[1] + fy) / RW), np.float32(0.5)))
    got = run_hard(oracle_lib, d, pts, technique="rbsm_filtered", si=float(si), kernel_order=3, penumbra_size=1,
                   max_search=16, depth_threshold=thr)
    for (x, y, z), g in zip(pts, got):
        acc = np.float32(0.0)
        for tx, ty in _rpcf_taps(x, y, RW):
            v = np.float32(oracle_lib.rbsm_tap(d, p, tx, ty, z, filtered=True)["value"])
            acc = np.float32(acc + np.float32(v * (np.float32(1.0) - si) + si))
        assert abs(g - acc / np.float32(16.0)) <= 1e-6
