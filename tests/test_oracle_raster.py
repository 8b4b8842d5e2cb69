"""Pins of the oracle's depth rasterizer and G-buffer (a2, a6).

The reference rasterizes with the OpenGL driver; its raster rules live outside
the repository (Mesa llvmpipe or the vendor driver), so these tests pin the
oracle's stated rules instead: coverage by pixel centre with a top-left tie
rule, FP32 plane depth, clipping, polygon offset, perspective-correct
attributes.  Expected values come from closed forms in float64 and from a
per-texel brute-force evaluation.
"""
import numpy as np
import pytest


def ortho_unit():
    """Column-major matrix mapping x, y in [0, 1] to NDC [-1, 1] and z in
    [0, 1] to NDC [-1, 1] (window depth = z)."""
    m = np.zeros((4, 4), np.float32)
    m[0, 0] = 2.0
    m[0, 3] = -1.0
    m[1, 1] = 2.0
    m[1, 3] = -1.0
    m[2, 2] = 2.0
    m[2, 3] = -1.0
    m[3, 3] = 1.0
    return np.ascontiguousarray(m.T).reshape(-1)


def quad(x0, x1, y0, y1, z, ccw=True):
    pos = np.array([[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]] if ccw else [[0, 2, 1], [0, 3, 2]], np.int32)
    return pos, idx


def test_bbox_raster_equals_brute_force(oracle_lib):
    rng = np.random.default_rng(0)
    for trial in range(6):
        n = 12
        pos = rng.uniform(-0.4, 1.4, size=(3 * n, 3)).astype(np.float32)
        pos[:, 2] = rng.uniform(-0.2, 1.2, size=3 * n)  # some cross near/far planes
        idx = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
        m = ortho_unit()
        for off in [(0.0, 0.0), (4.0, 20.0)]:
            a = oracle_lib.depth_raster(pos, idx, m, 24, 19, off)
            b = oracle_lib.depth_raster(pos, idx, m, 24, 19, off, brute=True)
            assert (a.view(np.uint32) == b.view(np.uint32)).all(), (trial, off)


def test_axis_aligned_quad_coverage_and_tie_rule(oracle_lib):
    """Edges exactly through pixel centres: the top-left rule (GL's y-up window
    space) includes the left and top edges and excludes the right and bottom."""
    W = H = 16
    # x in [2.5, 9.5] px, y in [3.5, 11.5] px (pixel centres i + 0.5)
    pos, idx = quad(2.5 / W, 9.5 / W, 3.5 / H, 11.5 / H, 0.25)
    for ccw in (True, False):
        p, t = quad(2.5 / W, 9.5 / W, 3.5 / H, 11.5 / H, 0.25, ccw=ccw)
        d = oracle_lib.depth_raster(p, t, ortho_unit(), W, H)
        cover = d < 1.0
        cx = np.arange(W) + 0.5
        cy = np.arange(H) + 0.5
        want = (cx[None, :] >= 2.5) & (cx[None, :] < 9.5) & (cy[:, None] > 3.5) & (cy[:, None] <= 11.5)
        np.testing.assert_array_equal(cover, want)
        np.testing.assert_array_equal(d[cover], np.float32(0.25))


def test_shared_edges_cover_each_pixel_once(oracle_lib):
    """A fan of triangles around an interior point tiles a polygon: every pixel
    centre inside is covered by exactly one triangle (no cracks, no double
    hits), whatever the edge slopes."""
    W = H = 40
    rng = np.random.default_rng(1)
    c = np.array([0.5, 0.5])
    ang = np.sort(rng.uniform(0, 2 * np.pi, 9))
    ring = c + 0.4 * np.stack([np.cos(ang), np.sin(ang)], 1)
    m = ortho_unit()
    hits = np.zeros((H, W), int)
    for k in range(len(ring)):
        tri = np.array([[*c, 0.5], [*ring[k], 0.5], [*ring[(k + 1) % len(ring)], 0.5]], np.float32)
        d = oracle_lib.depth_raster(tri, np.array([[0, 1, 2]], np.int32), m, W, H)
        hits += d < 1.0
    assert hits.max() == 1
    # union equals the float64 point-in-polygon test away from the boundary
    cx, cy = np.meshgrid((np.arange(W) + 0.5) / W, (np.arange(H) + 0.5) / H)
    inside = np.ones_like(cx, bool)
    margin = np.full_like(cx, np.inf)
    for k in range(len(ring)):
        a, b = ring[k], ring[(k + 1) % len(ring)]
        e = (b[0] - a[0]) * (cy - a[1]) - (b[1] - a[1]) * (cx - a[0])
        inside &= e > 0
        margin = np.minimum(margin, np.abs(e) / np.hypot(*(b - a)))
    far = margin * W > 0.01
    np.testing.assert_array_equal((hits == 1)[far], inside[far])


def test_plane_depth_is_linear(oracle_lib):
    """A slanted triangle under an orthographic view: window depth at each
    covered pixel centre equals the plane through the vertices (float64)
    within FP32 interpolation error."""
    W, H = 64, 48
    tri = np.array([[0.05, 0.1, 0.2], [0.95, 0.2, 0.8], [0.3, 0.9, 0.55]], np.float32)
    d = oracle_lib.depth_raster(tri, np.array([[0, 1, 2]], np.int32), ortho_unit(), W, H)
    cover = d < 1.0
    assert cover.sum() > 500
    # window coordinates as the raster computes them: FP32 viewport transform,
    # snapped to 1/256 pixel; the plane passes through the snapped vertices
    f32 = np.float32
    xn, yn, zn = (f32(2) * tri[:, 0] - f32(1)), (f32(2) * tri[:, 1] - f32(1)), (f32(2) * tri[:, 2] - f32(1))
    xw = np.rint((xn * f32(W / 2) + f32(W / 2)) * f32(256)) / 256.0
    yw = np.rint((yn * f32(H / 2) + f32(H / 2)) * f32(256)) / 256.0
    zw = (zn * f32(0.5) + f32(0.5)).astype(np.float64)
    P = np.c_[xw, yw, np.ones(3)]
    coef = np.linalg.solve(P, zw)
    cx, cy = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    want = coef[0] * cx + coef[1] * cy + coef[2]
    np.testing.assert_allclose(d[cover], want[cover], atol=3e-7)


def test_perspective_depth_matches_window_depth(oracle_lib, gi):
    """Perspective light: window depth is linear in screen space; at each
    pixel centre it equals the depth of the plane point seen there (float64
    ray-plane intersection pushed through P and the viewport)."""
    from globalillumination_amd import camera

    W = H = 96
    P = camera.perspective(45.0, 1.0, 1.0, 1000.0)
    V = camera.lookat((4.0, 8.0, 4.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0))
    mvp = camera.mat4_mul(P, V)
    tri = np.array([[-2.0, 0.0, -2.0], [2.5, 0.3, -1.0], [0.0, -0.2, 2.5]], np.float32)
    d = oracle_lib.depth_raster(tri, np.array([[0, 1, 2]], np.int32), mvp, W, H)
    cover = d < 1.0
    assert cover.sum() > 500
    M = np.asarray(mvp, np.float64).reshape(4, 4).T
    inv = np.linalg.inv(M)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0]).astype(np.float64)
    js, is_ = np.nonzero(cover)
    xn = (is_ + 0.5) / W * 2 - 1
    yn = (js + 0.5) / H * 2 - 1
    a = inv @ np.stack([xn, yn, -np.ones_like(xn), np.ones_like(xn)])
    b = inv @ np.stack([xn, yn, np.ones_like(xn), np.ones_like(xn)])
    a, b = a[:3] / a[3], b[:3] / b[3]
    t = ((tri[0].astype(np.float64)[:, None] - a) * n[:, None]).sum(0) / ((b - a) * n[:, None]).sum(0)
    X = a + t * (b - a)
    c = M @ np.vstack([X, np.ones(X.shape[1])])
    zw = c[2] / c[3] * 0.5 + 0.5
    np.testing.assert_allclose(d[js, is_], zw, atol=2e-6)


def test_raster_is_order_independent(oracle_lib):
    rng = np.random.default_rng(2)
    n = 40
    pos = rng.uniform(0.0, 1.0, size=(3 * n, 3)).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    a = oracle_lib.depth_raster(pos, idx, ortho_unit(), 50, 50)
    b = oracle_lib.depth_raster(pos, idx[rng.permutation(n)], ortho_unit(), 50, 50)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()
    assert a.max() <= 1.0 and a.min() >= 0.0


def test_clipping_keeps_viewport_coverage(oracle_lib, gi):
    """A ground quad reaching far beyond the view (guard band) and behind the
    near plane: coverage inside the viewport equals the float64 projection of
    the unclipped quad away from its edges."""
    from globalillumination_amd import camera

    W = H = 64
    P = camera.perspective(60.0, 1.0, 1.0, 100.0)
    V = camera.lookat((0.0, 2.0, 0.0), (0.0, 0.0, -6.0), (0.0, 1.0, 0.0))
    mvp = camera.mat4_mul(P, V)
    pos, idx = quad(-200.0, 200.0, -200.0, 200.0, 0.0)
    pos = pos[:, [0, 2, 1]].copy()  # the plane y = 0
    pos[:, 2] *= -1
    d = oracle_lib.depth_raster(pos, idx, mvp, W, H)
    M = np.asarray(mvp, np.float64).reshape(4, 4).T
    inv = np.linalg.inv(M)
    cx, cy = np.meshgrid((np.arange(W) + 0.5) / W * 2 - 1, (np.arange(H) + 0.5) / H * 2 - 1)
    a = inv @ np.stack([cx.ravel(), cy.ravel(), -np.ones(cx.size), np.ones(cx.size)])
    b = inv @ np.stack([cx.ravel(), cy.ravel(), np.ones(cx.size), np.ones(cx.size)])
    a, b = a[:3] / a[3], b[:3] / b[3]
    t = -a[1] / (b[1] - a[1])  # ray hits y = 0 between near and far
    hit = ((t >= 0) & (t <= 1)).reshape(H, W)
    X = a + t * (b - a)
    edge = (np.abs(t - 0) * 0 + np.minimum(np.abs(t), np.abs(1 - t))).reshape(H, W)
    ok = edge > 1e-3
    np.testing.assert_array_equal((d < 1.0)[ok], hit[ok])
    assert hit.mean() > 0.3


def test_polygon_offset_closed_form(oracle_lib):
    """glPolygonOffset(factor, units) on a float depth buffer: z + factor*m +
    units*2^(e-23), e the exponent of the largest |z| (m = 0 for a flat
    triangle)."""
    W = H = 8
    pos, idx = quad(0.0, 1.0, 0.0, 1.0, 0.5)
    d = oracle_lib.depth_raster(pos, idx, ortho_unit(), W, H, (4.0, 20.0))
    want = np.float32(np.float32(0.5) + np.float32(20.0 * 2.0 ** (-1 - 23)))
    np.testing.assert_array_equal(d, want)
    # slanted: dz/dx = 0.01 per pixel (window units), dz/dy = 0 -> factor * 0.01
    tri = np.array([[0.0, 0.0, 0.25], [1.0, 0.0, 0.25 + 0.01 * W], [0.0, 1.0, 0.25]], np.float32)
    base = oracle_lib.depth_raster(tri, np.array([[0, 1, 2]], np.int32), ortho_unit(), W, H)
    offd = oracle_lib.depth_raster(tri, np.array([[0, 1, 2]], np.int32), ortho_unit(), W, H, (4.0, 0.0))
    cov = base < 1.0
    assert cov.sum() > 20
    np.testing.assert_allclose(offd[cov] - base[cov], 0.04, atol=2e-7)


def test_gbuffer_tilted_plane(oracle_lib, gi):
    """Perspective-correct attributes: on a plane tilted away from the camera
    (w varies across it), the G-buffer position at each covered pixel is the
    float64 ray-plane hit, up to the 1/256-pixel vertex snap (a screen-linear
    interpolation would be off by ~1e-1 here); back-facing winding negates the
    normal; uncovered pixels stay 0 (the reference's background marker)."""
    from globalillumination_amd import camera

    W, H = 40, 30
    P = camera.perspective(60.0, W / H, 1.0, 100.0)
    V = camera.lookat((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    mvp = camera.mat4_mul(P, V)
    # plane z = 0.3 - 1.5 y
    pos = np.array([[-1.5, -1.0, 1.8], [2.0, -1.0, 1.8], [2.0, 1.2, -1.5], [-1.5, 1.2, -1.5]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    nrm = np.tile(np.array([[0.0, 0.83205, 0.5547]], np.float32), (4, 1))
    M = np.asarray(mvp, np.float64).reshape(4, 4).T
    inv = np.linalg.inv(M)
    px_world = 2 * 6.8 * np.tan(np.radians(30)) / H  # world size of a pixel at the far end
    for ccw, sign in [(True, 1.0), (False, -1.0)]:
        p_idx = idx if ccw else idx[:, [0, 2, 1]]
        gp, gn, _ = oracle_lib.gbuffer(pos, nrm, p_idx, mvp, W, H)
        cover = gp[0] != 0
        assert cover.sum() > 100
        js, is_ = np.nonzero(cover)
        xn, yn = (is_ + 0.5) / W * 2 - 1, (js + 0.5) / H * 2 - 1
        a = inv @ np.stack([xn, yn, -np.ones_like(xn), np.ones_like(xn)])
        b = inv @ np.stack([xn, yn, np.ones_like(xn), np.ones_like(xn)])
        a, b = a[:3] / a[3], b[:3] / b[3]
        n = np.array([0.0, 1.5, 1.0])
        t = (0.3 - (a * n[:, None]).sum(0)) / (((b - a) * n[:, None]).sum(0))
        X = a + t * (b - a)
        tol = 4 * px_world / 256
        for c in range(3):
            np.testing.assert_allclose(gp[c][cover], X[c], atol=tol)
        np.testing.assert_allclose(gn[1][cover], sign * 0.83205, atol=1e-6)
        assert (gn[:, ~cover] == 0).all() and (gp[:, ~cover] == 0).all()
