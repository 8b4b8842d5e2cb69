"""Pins of the oracle's PCSS (PlausibleSoftShadow.frag with PCSS == 1):
closed forms on half-plane and uniform maps, and the reference's quirks
(the penumbra cut-off at average blocker depth 0.99, the 1024-texel search
rule, fractional loop bounds, zNear left at 0)."""
import numpy as np
import pytest

SI = 0.25


def _run(oracle_lib, depth, pts, zn=1, bss=7, ks=15, radius=8, si=SI):
    pts = np.asarray(pts, np.float32)
    pos = np.ascontiguousarray(pts.T.reshape(3, 1, -1))
    h, w = depth.shape
    p = oracle_lib.make_params(np.eye(4, dtype=np.float32).reshape(-1), w, h, "naive", shadow_intensity=si)
    return oracle_lib.pcss(pos, None, depth, p, blocker_search_size=bss, kernel_size=ks, light_source_radius=radius,
                           z_near=zn)[0]


def _half_plane(n=1024, edge=512, blocker=0.995):
    d = np.ones((n, n), np.float32)
    d[:, :edge] = blocker
    return d


def test_pcss_half_plane_closed_form(oracle_lib):
    """A pixel on the edge of a half-plane occluder at depth 0.995 (receiver
    0.999, radius 8, zNear 1): the blocker search finds the occluder, the
    penumbra is ~0.032 in texture space, and of the 15 x 15 PCF taps the 7
    columns with w < 0 fall on the occluder: (8*15 + 7*15*si) / 225.  Deep in
    the occluder every tap is shadowed (si); on the lit side far from the edge
    no blocker is found and the result is 1."""
    d = _half_plane()
    got = _run(oracle_lib, d, [(0.5, 0.5, 0.999), (0.2, 0.5, 0.999), (0.8, 0.5, 0.999)])
    np.testing.assert_allclose(got, [(120 + 105 * SI) / 225, SI, 1.0], rtol=1e-6)


def test_pcss_even_kernel_runs_fewer_taps_than_it_divides_by(oracle_lib):
    """kernelSize 4: filterWidth 1.5, the loop runs h = int(-1.5) = -1 .. 1,
    3 x 3 taps, but divides by 16: a fully shadowed pixel gives 9 si / 16."""
    d = _half_plane()
    got = _run(oracle_lib, d, [(0.2, 0.5, 0.999)], ks=4)
    np.testing.assert_allclose(got, [9 * SI / 16], rtol=1e-6)
    # blockerSearchSize 4 (3 x 3 search taps) still finds the occluder
    got = _run(oracle_lib, d, [(0.2, 0.5, 0.999)], bss=4)
    np.testing.assert_allclose(got, [SI], rtol=1e-6)


def test_pcss_shallow_blockers_give_no_penumbra(oracle_lib):
    """computePenumbraWidth returns 0 when the average blocker depth is below
    0.99; PCF then returns 1, so a pixel fully behind a near occluder is lit."""
    d = _half_plane(blocker=0.5)
    got = _run(oracle_lib, d, [(0.2, 0.5, 0.999), (0.5, 0.5, 0.999)])
    np.testing.assert_array_equal(got, [1.0, 1.0])


def test_pcss_znear_zero_collapses_to_lit(oracle_lib):
    """The reference never uploads zNear for this shader, so it is 0 and the
    penumbra width is 0 * w = 0: every pixel the pre-evaluation lets through
    is lit, whatever the map holds."""
    d = _half_plane()
    rng = np.random.default_rng(0)
    pts = np.column_stack([rng.uniform(0.05, 0.95, 200), rng.uniform(0.05, 0.95, 200),
                           rng.uniform(0.9, 1.0, 200)]).astype(np.float32)
    np.testing.assert_array_equal(_run(oracle_lib, d, pts, zn=0), np.ones(200, np.float32))


def test_pcss_search_width_uses_1024_for_large_maps(oracle_lib):
    """blockerSearchWidth = radius / min(width, 1024): on a 2048 map the
    outer search taps sit 16 texels away, not 8.  A blocker band 9-12 texels
    to the left is found only by the 16-texel rule (its tap at -10.67 texels;
    the 8-texel rule's taps sit at -8, -5.33, -2.67), and then the PCF column
    at -9.4 texels (penumbra 0.0322, 66 texels, over 7 steps) falls on it:
    15 of 225 taps shadowed."""
    n = 2048
    d = np.ones((n, n), np.float32)
    d[:, 1024 - 12:1024 - 8] = 0.995
    x = np.float32((1024 + 0.5) / n)
    got = _run(oracle_lib, d, [(x, 0.5, 0.999)])
    np.testing.assert_allclose(got, [(210 + 15 * SI) / 225], rtol=1e-6)
    # the same band moved to 4-6 texels left, inside the first ring of both
    # rules but between the 16-texel rule's taps at -5.33 and 0: not found
    d2 = np.ones((n, n), np.float32)
    d2[:, 1024 - 4:1024 - 2] = 0.995
    assert _run(oracle_lib, d2, [(x, 0.5, 0.999)])[0] == 1.0


def test_pcss_uniform_maps(oracle_lib):
    rng = np.random.default_rng(1)
    pts = np.column_stack([rng.uniform(0.1, 0.9, 100), rng.uniform(0.1, 0.9, 100),
                           rng.uniform(0.995, 0.999, 100)]).astype(np.float32)
    # nothing in front: lit
    np.testing.assert_array_equal(_run(oracle_lib, np.ones((256, 256), np.float32), pts), np.ones(100))
    # everything in front at 0.99: umbra everywhere (si), since the penumbra is positive and small
    got = _run(oracle_lib, np.full((256, 256), 0.99, np.float32), pts)
    np.testing.assert_allclose(got, np.full(100, SI), rtol=1e-6)


def test_pcss_rejects_unit_sizes(oracle_lib):
    d = np.ones((16, 16), np.float32)
    with pytest.raises(RuntimeError):
        _run(oracle_lib, d, [(0.5, 0.5, 0.5)], ks=1)
    with pytest.raises(RuntimeError):
        _run(oracle_lib, d, [(0.5, 0.5, 0.5)], bss=1)
