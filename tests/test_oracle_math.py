"""Pins of the oracle's light-camera math (a1, a3, a4) against closed forms.

The reference ships no tests or golden vectors for this path, so the oracle's
glm restatement is checked against the closed-form matrices in float64, and a
plausible mistake (radians instead of degrees, a transposed matrix, a wrong
sign, a bias matrix off by one column) fails one of these checks.
"""
import math

import numpy as np
import pytest

RTOL = 4e-7  # a few float32 ulps


def as_rows(m16):
    """Column-major 16 floats -> row-major 4x4 float64."""
    return np.asarray(m16, dtype=np.float64).reshape(4, 4).T


@pytest.mark.parametrize("fovy,aspect,zn,zf", [(45.0, 1.0, 1.0, 1000.0), (60.0, 800 / 600, 1.0, 1000.0),
                                               (30.0, 16 / 9, 0.5, 50.0)])
def test_perspective_closed_form(oracle_lib, fovy, aspect, zn, zf):
    m = as_rows(oracle_lib.perspective(fovy, aspect, zn, zf))
    t = math.tan(math.radians(fovy) / 2.0)  # fovy is in DEGREES (glm 0.9.3)
    want = np.zeros((4, 4))
    want[0, 0] = 1.0 / (aspect * t)
    want[1, 1] = 1.0 / t
    want[2, 2] = -(zf + zn) / (zf - zn)
    want[2, 3] = -2.0 * zf * zn / (zf - zn)
    want[3, 2] = -1.0
    np.testing.assert_allclose(m, want, rtol=RTOL, atol=1e-30)


def test_ortho_closed_form(oracle_lib):
    l, r, b, t, n, f = -3.0, 5.0, -2.0, 6.0, 1.5, 40.0
    m = as_rows(oracle_lib.ortho(l, r, b, t, n, f))
    want = np.array([[2 / (r - l), 0, 0, -(r + l) / (r - l)],
                     [0, 2 / (t - b), 0, -(t + b) / (t - b)],
                     [0, 0, -2 / (f - n), -(f + n) / (f - n)],
                     [0, 0, 0, 1.0]])
    np.testing.assert_allclose(m, want, rtol=RTOL, atol=1e-7)


@pytest.mark.parametrize("eye,at", [((4.0, 8.0, 4.0), (0.0, 0.0, 0.0)), ((0.0, 3.0, 6.0), (0.0, 1.0, 0.0)),
                                    ((-16.3, 56.4, -10.9), (0.0, 2.08, 0.0))])
def test_lookat_maps_eye_and_center(oracle_lib, eye, at):
    up = (0.0, 0.0, 1.0)
    V = as_rows(oracle_lib.lookat(eye, at, up))
    R = V[:3, :3]
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=2e-6)  # rotation
    assert np.linalg.det(R) > 0
    e = V @ np.array([*eye, 1.0])
    np.testing.assert_allclose(e[:3], 0.0, atol=2e-5)  # eye -> origin
    c = V @ np.array([*at, 1.0])
    d = np.linalg.norm(np.subtract(at, eye))
    np.testing.assert_allclose(c[:3], [0.0, 0.0, -d], atol=5e-5 * max(1.0, d))  # looks down -z
    u = R @ np.array(up)
    assert u[1] > 0 and abs(u[0]) < 1e-5  # up vector lands in the y-z plane, pointing +y


def test_mat4_mul_matches_float64(oracle_lib):
    rng = np.random.default_rng(3)
    a = rng.normal(size=16).astype(np.float32)
    b = rng.normal(size=16).astype(np.float32)
    got = as_rows(oracle_lib.mat4_mul(a, b))
    want = as_rows(a) @ as_rows(b)  # glm: (A*B) with column-major storage
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("zbias", [0.0, 0.005])
def test_bias_matrix_maps_ndc_to_texture_space(oracle_lib, zbias):
    B = as_rows(oracle_lib.bias_matrix(zbias))
    for p in [(-1, -1, -1), (1, 1, 1), (0.25, -0.5, 0.75)]:
        q = B @ np.array([*p, 1.0])
        np.testing.assert_allclose(q[:3], [0.5 * p[0] + 0.5, 0.5 * p[1] + 0.5, 0.5 * p[2] + 0.5 - zbias],
                                   atol=1e-7)
        assert q[3] == 1.0


def test_uniform_sampling_grid(oracle_lib):
    """UniformSampledLightSource: a sqrt(N) x sqrt(N) grid of side `size` in
    world XY, row-major from the (-x, -y) corner (N = 289, size = 16 as in
    SoftShadowMapping/src/main.cpp:1598-1600)."""
    base = (10.0, 130.0, 100.0)
    pts = np.array([oracle_lib.uniform_sampling(base, 16, 289, i) for i in range(289)], dtype=np.float64)
    k = np.arange(289)
    want_x = base[0] + ((k % 17) - 8) / 8.0 * 8.0
    want_y = base[1] + ((k // 17) - 8) / 8.0 * 8.0
    np.testing.assert_allclose(pts[:, 0], want_x, atol=1e-4)
    np.testing.assert_allclose(pts[:, 1], want_y, atol=1e-4)
    np.testing.assert_allclose(pts[:, 2], base[2])


def test_host_math_matches_oracle_bits(oracle_lib, gi):
    """The C ABI's host math (product) and the oracle restate the same glm
    formulas in FP32 and must agree bit for bit."""
    from globalillumination_amd import camera

    pairs = [
        (camera.perspective(45.0, 1.0, 1.0, 1000.0), oracle_lib.perspective(45.0, 1.0, 1.0, 1000.0)),
        (camera.perspective(60.0, 1920 / 1080, 1.0, 1000.0), oracle_lib.perspective(60.0, 1920 / 1080, 1.0, 1000.0)),
        (camera.ortho(-3, 5, -2, 6, 1.5, 40), oracle_lib.ortho(-3, 5, -2, 6, 1.5, 40)),
        (camera.lookat((4, 8, 4), (0, 0, 0), (0, 0, 1)), oracle_lib.lookat((4, 8, 4), (0, 0, 0), (0, 0, 1))),
        (camera.bias_matrix(0.005), oracle_lib.bias_matrix(0.005)),
        (camera.uniform_sampling((1, 2, 3), 16, 289, 37), oracle_lib.uniform_sampling((1, 2, 3), 16, 289, 37)),
    ]
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=16).astype(np.float32), rng.normal(size=16).astype(np.float32)
    pairs.append((camera.mat4_mul(a, b), oracle_lib.mat4_mul(a, b)))
    for got, want in pairs:
        assert (np.asarray(got).view(np.uint32) == np.asarray(want).view(np.uint32)).all()


def test_soft_light_trans_is_column_three(gi):
    from globalillumination_amd import camera

    rng = np.random.default_rng(9)
    m = rng.normal(size=(5, 16)).astype(np.float32)
    t = camera.soft_light_trans(m)
    np.testing.assert_array_equal(t, m[:, 12:16])
