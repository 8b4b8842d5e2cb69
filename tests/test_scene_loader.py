"""The scene configuration / OBJ loader (include/gi_scene.h) on synthetic
files: glmReadOBJ's face formats, fans and negative indices, computeNormals'
running mean, the s/t/r/c/cf/m/+ lines in file order, and the camera, light
and depth-threshold lines (ShadowMapping/src/IO/SceneLoader.cpp:12-91,
ShadowMapping/src/Mesh.cpp).  When the reference tree is present, its own
configurations load with the triangle counts its scenes are quoted at."""
import math
import os
from pathlib import Path

import numpy as np
import pytest


@pytest.fixture(scope="module")
def scenes():
    from globalillumination_amd import scenes as s

    return s


def _write(p: Path, text: str) -> Path:
    p.write_text(text)
    return p


def test_obj_face_formats_fans_and_negative_indices(tmp_path, scenes):
    obj = _write(tmp_path / "m.obj", """# comment
mtllib x.mtl
o thing
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vn 0 0 1
g group one
usemtl red
s off
f 1 2 3
f 1/1 3/3 4/1
f 1//1 2//1 4//1
f 2/2/1 3/3/1 4/3/1
f -4 -3 -2 -1
""")
    m = scenes.load_obj(obj)
    assert m.nverts == 4
    want = [[0, 1, 2], [0, 2, 3], [0, 1, 3], [1, 2, 3], [0, 1, 2], [0, 2, 3]]  # the quad is a fan
    np.testing.assert_array_equal(m.indices, want)


def test_obj_rejects_out_of_range_index_and_missing_file(tmp_path, scenes):
    from globalillumination_amd import _lib

    obj = _write(tmp_path / "bad.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(_lib.GiError) as e:
        scenes.load_obj(obj)
    assert e.value.status == _lib.GI_EIO
    with pytest.raises(_lib.GiError):
        scenes.load_obj(tmp_path / "missing.obj")


def test_compute_normals_face_and_running_mean(tmp_path, scenes):
    """Face normal = normalize(cross(b - a, c - a)); a vertex shared by k faces
    holds the plain mean of their unit normals (not renormalised): here
    (0, 0, 1) and (1, 0, 0) give (0.5, 0, 0.5) at the shared edge."""
    obj = _write(tmp_path / "fold.obj", """v 0 0 0
v 0 1 0
v 1 0 0
v 0 0 -1
f 1 3 2
f 1 4 2
""")
    m = scenes.load_obj(obj)
    np.testing.assert_allclose(m.normals[2], [0, 0, 1], atol=0)  # cross((1,0,0), (0,1,0))
    np.testing.assert_allclose(m.normals[3], [1, 0, 0], atol=0)  # cross((0,0,-1), (0,1,0))
    np.testing.assert_allclose(m.normals[0], [0.5, 0, 0.5], atol=1e-7)
    np.testing.assert_allclose(m.normals[1], [0.5, 0, 0.5], atol=1e-7)


def test_compute_normals_mean_of_three(tmp_path, scenes):
    """Three faces around a corner with normals x, y, z: the running mean
    n*(1 - 1/k) + f/k gives (1/3, 1/3, 1/3)."""
    obj = _write(tmp_path / "corner.obj", """v 0 0 0
v 1 0 0
v 0 1 0
v 0 0 1
f 1 2 3
f 1 3 4
f 1 4 2
""")
    m = scenes.load_obj(obj)
    np.testing.assert_allclose(m.normals[0], [1 / 3, 1 / 3, 1 / 3], atol=2e-7)


def test_config_transforms_in_file_order(tmp_path, scenes):
    """s, t and r apply to positions in file order; only r turns normals
    (Mesh::scale/translate leave them); glm's rotate takes degrees, and
    rotate(90, x) maps +y to +z.  `+` appends with an index offset."""
    _write(tmp_path / "tri.obj", "v 0 1 0\nv 1 1 0\nv 0 1 1\nf 1 3 2\n")
    cfg = _write(tmp_path / "scene.txt", """o tri.obj
s 2.0 2.0 2.0
t 1.0 0.0 0.0
c 0.5 0.25 1.0
+
o tri.obj
r 90.0 0.0 0.0
+
ve 0.0 41.0 -50.0
va 0.0 16.0 -10.0
le 10.0 130.0 100.0
la 1.0 2.0 3.0
d 0.000025
ha 0.5
hb 0.001
""")
    wl = scenes.load_scene(cfg, base_dir=tmp_path)
    m = wl.mesh
    assert m.nverts == 6 and m.ntris == 2
    np.testing.assert_array_equal(m.indices, [[0, 2, 1], [3, 5, 4]])
    np.testing.assert_allclose(m.positions[:3], [[1, 2, 0], [3, 2, 0], [1, 2, 2]], atol=0)
    # normal of (0,1,0),(0,1,1),(1,1,0): cross((0,0,1),(1,0,0)) = (0,1,0); scale leaves it
    np.testing.assert_allclose(m.normals[:3], np.tile([0, 1, 0], (3, 1)), atol=0)
    # rotated 90 degrees about x: y -> z, z -> -y
    np.testing.assert_allclose(m.positions[3:], [[0, 0, 1], [1, 0, 1], [0, -1, 1]], atol=1e-6)
    np.testing.assert_allclose(m.normals[3:], np.tile([0, 0, 1], (3, 1)), atol=1e-6)
    assert wl.camera.eye == pytest.approx((0.0, 41.0, -50.0)) and wl.camera.at == pytest.approx((0.0, 16.0, -10.0))
    assert wl.lights[0].eye == pytest.approx((10.0, 130.0, 100.0)) and wl.lights[0].at == pytest.approx((1, 2, 3))
    assert wl.camera.up == (0.0, 0.0, 1.0) and wl.camera.fov == 45.0
    assert wl.extra["depth_threshold"] == pytest.approx(0.000025)
    assert wl.extra["hsm"] == pytest.approx((0.5, 0.001))
    # colours: 3 vertices of the first object; the second added none (zero-padded here)
    np.testing.assert_array_equal(wl.extra["colors"][:3], np.tile(np.float32([0.5, 0.25, 1.0]), (3, 1)))
    np.testing.assert_array_equal(wl.extra["colors"][3:], 0)


def test_config_rotation_order_and_colour_file(tmp_path, scenes):
    """r x y z is R = Rx * Ry * Rz applied to column vectors: (1, 0, 0) under
    r 90 0 90 goes Rz first to (0, 1, 0), then Rx to (0, 0, 1).  `cf` reads
    per-vertex colours from `v x y z r g b` lines; `m` records texture paths."""
    _write(tmp_path / "p.obj", "v 1 0 0 0.1 0.2 0.3\nv 2 0 0 0.4 0.5 0.6\nv 1 1 0 0.7 0.8 0.9\nf 1 2 3\n")
    cfg = _write(tmp_path / "s.txt", "o p.obj\nr 90 0 90\ncf p.obj\nm tex/white.png\n+\n")
    wl = scenes.load_scene(cfg, base_dir=tmp_path)
    np.testing.assert_allclose(wl.mesh.positions[0], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(wl.extra["colors"], [[0.1, 0.2, 0.3], [0.4, 0.5, 0.6], [0.7, 0.8, 0.9]], rtol=1e-7)
    assert wl.extra["textures"] == ["tex/white.png"]


def test_config_errors(tmp_path, scenes):
    from globalillumination_amd import _lib

    with pytest.raises(_lib.GiError):
        scenes.load_scene(tmp_path / "none.txt")
    cfg = _write(tmp_path / "bad.txt", "s 1 1 1\n")  # a transform before any object
    with pytest.raises(_lib.GiError):
        scenes.load_scene(cfg)


REF = Path("/root/reference")


def _fan_triangles(obj: Path) -> tuple:
    """(vertices, triangles) of an OBJ file counted independently of the
    loader: one triangle per face corner after the second."""
    nv = nt = 0
    with open(obj, errors="replace") as f:
        for ln in f:
            parts = ln.split()
            if not parts:
                continue
            if parts[0] == "v":
                nv += 1
            elif parts[0] == "f":
                nt += len(parts) - 3
    return nv, nt


@pytest.mark.skipif(not (REF / "Configs").is_dir(), reason="the reference tree is not present")
@pytest.mark.parametrize("name", ["Teapot", "Raptor", "SanDiego", "Dragon", "TreeWithoutLeaves", "Door", "QuadBot"])
def test_reference_configurations_load(scenes, name):
    """The reference's own scenes load: vertex and triangle counts equal the
    sums over the OBJ files each configuration names, indices stay in range,
    positions are finite, and the depth threshold is read."""
    cfg = REF / "Configs" / f"{name}.txt"
    need = [ln.split()[1] for ln in cfg.read_text().splitlines() if ln.startswith("o ")]
    if not all((REF / p).exists() for p in need):
        pytest.skip(f"{name}: an OBJ file it names is not in the tree")
    counts = [_fan_triangles(REF / p) for p in need]
    wl = scenes.load_scene(cfg, base_dir=REF)
    assert wl.mesh.nverts == sum(c[0] for c in counts)
    assert wl.mesh.ntris == sum(c[1] for c in counts)
    assert np.isfinite(wl.mesh.positions).all()
    assert wl.mesh.indices.min() >= 0 and wl.mesh.indices.max() < wl.mesh.nverts
    assert wl.extra["depth_threshold"] > 0
