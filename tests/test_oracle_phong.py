"""Pins of the oracle's deferred Phong shading (PhongShading.frag:11-46)
against closed forms and against the invariants the shader fixes."""
import math

import numpy as np
import pytest

EYE3 = np.eye(3, dtype=np.float32).T.reshape(9)
EYE4 = np.eye(4, dtype=np.float32).T.reshape(16)


def _translate(t):
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = t
    return m.T.reshape(16)  # column-major


def _gbuf(verts, normals, facing):
    """(3, 1, n) planes from per-pixel vertices / normals."""
    v = np.asarray(verts, np.float32).T.reshape(3, 1, -1)
    n = np.asarray(normals, np.float32).T.reshape(3, 1, -1)
    f = np.asarray(facing, np.float32).reshape(1, -1)
    return v, n, f


def test_phong_closed_form_head_on(oracle_lib):
    """MV translates the world vertex (1, 0, 0) to (0, 0, -5) in eye space
    (a world x of 0 would be background); normal +z, light at the eye:
    L = N = R = +z, E = normalize((0, 0, 5, -1)) keeps the w component, so
    dot(R, E) = 5/sqrt(26); 0.3 * shininess is 3."""
    v, n, f = _gbuf([(1.0, 0.0, 0.0)], [(0.0, 0.0, 1.0)], [1.0])
    sh = np.ones((1, 1), np.float32)
    out = oracle_lib.phong(v, n, f, sh, _translate((-1.0, 0.0, -5.0)), EYE3, (0.0, 0.0, 0.0), 0.25,
                           color=(1, 1, 1, 1))[0, 0]
    s = (5.0 / math.sqrt(26.0)) ** 3
    want_rgb = 0.5 + 0.75 * 0.25 * s + 0.4
    want_a = 1.0 + 0.75 * 1.0 * s + 1.0
    np.testing.assert_allclose(out, [want_rgb, want_rgb, want_rgb, want_a], rtol=2e-7)


def test_phong_closed_form_oblique(oracle_lib):
    """Normal +z at the origin of eye space shifted to z = -4, light at
    (3, 0, 0) in eye space: L = (3, 0, 4)/5, diffuse 0.8; R = (-3, 0, 4)/5;
    E = (0, 0, 4, -1)/sqrt(17); spec = (16/(5 sqrt 17))^3."""
    v, n, f = _gbuf([(2.0, 0.0, 0.0)], [(0.0, 0.0, 1.0)], [1.0])
    sh = np.full((1, 1), 0.5, np.float32)
    col = (0.2, 0.4, 0.6, 1.0)
    out = oracle_lib.phong(v, n, f, sh, _translate((-2.0, 0.0, -4.0)), EYE3, (3.0, 0.0, 0.0), 0.25,
                           color=col)[0, 0]
    spec = (16.0 / (5.0 * math.sqrt(17.0))) ** 3
    ldif, lspe, amb = (0.5, 0.5, 0.5, 1.0), (0.25, 0.25, 0.25, 1.0), (0.4, 0.4, 0.4, 1.0)
    want = [0.5 * col[c] * (ldif[c] * 0.8 + 0.25 * lspe[c] * spec + amb[c]) for c in range(4)]
    np.testing.assert_allclose(out, want, rtol=3e-7)


def test_phong_modelview_and_normal_matrix(oracle_lib):
    """A world-space translate + rotation: MV moves the vertex into eye space
    and normalMatrix rotates the normal; the result equals the same pixel
    given directly in eye space with identity matrices.  A transposed MV or
    normal matrix fails it (the rotation is not symmetric)."""
    c, s_ = math.cos(0.7), math.sin(0.7)
    R = np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]])
    t = np.array([0.5, -1.0, -6.0])
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    mv = M.T.reshape(16).astype(np.float32)  # column-major
    nm = np.linalg.inv(R).T.T.reshape(9).astype(np.float32)
    pw = np.array([0.3, 0.2, 0.1])
    nw = np.array([0.0, 0.6, 0.8])
    light = (1.0, 2.0, 0.5)
    v, n, f = _gbuf([pw], [nw], [1.0])
    sh = np.full((1, 1), 0.8, np.float32)
    got = oracle_lib.phong(v, n, f, sh, mv, nm, light, 0.25, color=(1, 1, 1, 1))[0, 0]
    pe = R @ pw + t
    ne = R @ nw
    v2, n2, f2 = _gbuf([pe], [ne], [1.0])
    want = oracle_lib.phong(v2, n2, f2, sh, EYE4, EYE3, light, 0.25, color=(1, 1, 1, 1))[0, 0]
    np.testing.assert_allclose(got, want, rtol=1e-5)
    # and a transposed model-view gives something else
    bad = oracle_lib.phong(v, n, f, sh, M.reshape(16).astype(np.float32), nm, light, 0.25, color=(1, 1, 1, 1))
    assert np.abs(bad[0, 0] - want).max() > 1e-3


def test_phong_back_face_reads_raw_normal(oracle_lib):
    """The G-buffer stores back-face normals flipped with facing 0; shading
    must see the raw normal (GBuffer.frag:36 stores it unflipped), so a
    back-facing pixel with a flipped stored normal shades exactly like a
    front-facing pixel with the raw normal."""
    rng = np.random.default_rng(3)
    verts = rng.uniform(-1, 1, (64, 3)) + np.array([0, 0, -6.0])
    nrm = rng.normal(size=(64, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    sh = rng.uniform(0.25, 1.0, (1, 64)).astype(np.float32)
    v, n, f1 = _gbuf(verts, nrm, np.ones(64))
    _, nflip, f0 = _gbuf(verts, -nrm, np.zeros(64))
    a = oracle_lib.phong(v, n, f1, sh, EYE4, EYE3, (2.0, 3.0, 1.0), 0.25)
    b = oracle_lib.phong(v, nflip, f0, sh, EYE4, EYE3, (2.0, 3.0, 1.0), 0.25)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()


def test_phong_unlit_side_and_background(oracle_lib):
    """The light straight behind a surface that faces the eye: no diffuse,
    and with L antiparallel to N, -reflect(L, N) = L points away from the eye,
    so no specular either: rgba = shadow * color * ambient.  A pixel with
    vertex.x == 0 is discarded and keeps the clear colour."""
    v, n, f = _gbuf([(0.5, 0.0, -5.0), (0.0, 1.0, -3.0)], [(0.0, 0.0, 1.0), (0.0, 0.0, 1.0)], [1.0, 1.0])
    sh = np.array([[0.6, 1.0]], np.float32)
    col = (0.3, 0.5, 0.7, 0.9)
    clear = (0.63, 0.82, 0.96, 1.0)
    out = oracle_lib.phong(v, n, f, sh, EYE4, EYE3, (0.5, 0.0, -20.0), 0.25, color=col, clear_color=clear)
    amb = (0.4, 0.4, 0.4, 1.0)
    np.testing.assert_allclose(out[0, 0], [0.6 * col[c] * amb[c] for c in range(4)], rtol=1e-6)
    np.testing.assert_array_equal(out[0, 1], np.float32(clear))


def test_phong_specular_vanishes_at_shadow_intensity(oracle_lib):
    """specShadow = shadow - shadowIntensity: a pixel whose shadow value equals
    shadowIntensity has no specular term, so its value is the diffuse and
    ambient terms only, scaled by the shadow."""
    rng = np.random.default_rng(5)
    verts = rng.uniform(-1, 1, (32, 3)) + np.array([0, 0, -5.0])
    nrm = np.tile([0.0, 0.0, 1.0], (32, 1))
    v, n, f = _gbuf(verts, nrm, np.ones(32))
    sh = np.full((1, 32), 0.25, np.float32)
    light = np.array([0.2, 0.1, 2.0])
    out = oracle_lib.phong(v, n, f, sh, EYE4, EYE3, light, 0.25, color=(1, 1, 1, 1))[0]
    L = light - verts
    L /= np.linalg.norm(L, axis=1, keepdims=True)
    d = np.maximum(L[:, 2], 0.0)
    np.testing.assert_allclose(out[:, 0], 0.25 * (0.5 * d + 0.4), rtol=1e-6)
    np.testing.assert_allclose(out[:, 3], 0.25 * (1.0 * d + 1.0), rtol=1e-6)


def test_phong_color_plane_matches_uniform_color(oracle_lib):
    rng = np.random.default_rng(7)
    verts = rng.uniform(-1, 1, (16, 3)) + np.array([0, 0, -5.0])
    nrm = rng.normal(size=(16, 3))
    v, n, f = _gbuf(verts, nrm, (rng.uniform(size=16) > 0.5).astype(np.float32))
    sh = rng.uniform(0.25, 1.0, (1, 16)).astype(np.float32)
    col = (0.1, 0.2, 0.3, 0.4)
    a = oracle_lib.phong(v, n, f, sh, EYE4, EYE3, (1.0, 1.0, 1.0), 0.25, color=col)
    plane = np.broadcast_to(np.float32(col), (1, 16, 4))
    b = oracle_lib.phong(v, n, f, sh, EYE4, EYE3, (1.0, 1.0, 1.0), 0.25, color=(9, 9, 9, 9), color_plane=plane)
    assert (a.view(np.uint32) == b.view(np.uint32)).all()


def test_gbuffer_facing_plane(oracle_lib):
    """Facing is 1 where the counter-clockwise (front) side of the triangle
    is seen, 0 for the clockwise side and for background."""
    from globalillumination_amd import camera

    pos = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    nrm = np.tile(np.float32([0, 0, 1]), (3, 1))
    mvp_front = camera.mat4_mul(camera.perspective(45.0, 1.0, 1.0, 100.0), camera.lookat((0, 0, 5), (0, 0, 0),
                                                                                        (0, 1, 0)))
    mvp_back = camera.mat4_mul(camera.perspective(45.0, 1.0, 1.0, 100.0), camera.lookat((0, 0, -5), (0, 0, 0),
                                                                                       (0, 1, 0)))
    for mvp, want, sign in ((mvp_front, 1.0, 1.0), (mvp_back, 0.0, -1.0)):
        p, n, fac = oracle_lib.gbuffer(pos, nrm, np.array([[0, 1, 2]], np.int32), mvp, 32, 32)
        hit = p[0] != 0
        assert hit.sum() > 50
        assert (fac[hit] == want).all() and (fac[~hit] == 0).all()
        assert (n[2][hit] == sign).all()
