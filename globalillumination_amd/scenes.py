"""Deterministic synthetic scenes for the five shadow-pipeline workloads.

The reference loads OBJ meshes through its scene config files
(ShadowMapping/src/IO/SceneLoader.cpp:29-91); the workloads here are generated
instead, from fixed seeds with numpy's PCG64, so every run and every machine
sees the same triangles.  World space is y-up like the reference scenes
(Configs/*.txt put the ground plane at y = -8 under the objects), and every
light/camera uses the reference's hard-coded up vector (0, 0, 1)
(ShadowMapping/src/main.cpp:866-867).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

UP = (0.0, 0.0, 1.0)


@dataclass
class Mesh:
    positions: np.ndarray  # (nv, 3) float32
    normals: np.ndarray    # (nv, 3) float32
    indices: np.ndarray    # (nt, 3) int32

    @property
    def ntris(self) -> int:
        return int(self.indices.shape[0])

    @property
    def nverts(self) -> int:
        return int(self.positions.shape[0])

    @staticmethod
    def concat(*meshes: "Mesh") -> "Mesh":
        pos, nrm, idx, base = [], [], [], 0
        for m in meshes:
            pos.append(m.positions)
            nrm.append(m.normals)
            idx.append(m.indices + base)
            base += m.nverts
        return Mesh(np.ascontiguousarray(np.concatenate(pos), dtype=np.float32),
                    np.ascontiguousarray(np.concatenate(nrm), dtype=np.float32),
                    np.ascontiguousarray(np.concatenate(idx), dtype=np.int32))

    def bounds(self):
        return self.positions.min(axis=0), self.positions.max(axis=0)


def orient_outward(pos: np.ndarray, tris: np.ndarray, center) -> np.ndarray:
    """Reorders each triangle so it winds CCW seen from outside `center`."""
    p = pos.astype(np.float64)
    fn = np.cross(p[tris[:, 1]] - p[tris[:, 0]], p[tris[:, 2]] - p[tris[:, 0]])
    out = np.einsum("ij,ij->i", fn, p[tris].mean(axis=1) - np.asarray(center, dtype=np.float64)[None]) > 0
    return np.where(out[:, None], tris, tris[:, [0, 2, 1]])


def uv_sphere(radius: float, center, segments: int = 64, rings: int = 32) -> Mesh:
    """UV sphere: `segments` around, `rings` bands (64 x 32 -> 3968 triangles)."""
    c = np.asarray(center, dtype=np.float64)
    theta = np.linspace(0.0, np.pi, rings + 1)[1:-1]
    phi = np.linspace(0.0, 2.0 * np.pi, segments, endpoint=False)
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    ring = np.stack([st * np.cos(phi), np.repeat(ct, segments, axis=1), st * np.sin(phi)], axis=-1).reshape(-1, 3)
    dirs = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]])
    top, bot = 0, dirs.shape[0] - 1
    idx = []
    s = segments
    for k in range(s):
        idx.append((top, 1 + (k + 1) % s, 1 + k))
    for r in range(rings - 2):
        a0, b0 = 1 + r * s, 1 + (r + 1) * s
        for k in range(s):
            k1 = (k + 1) % s
            idx.append((a0 + k, a0 + k1, b0 + k))
            idx.append((a0 + k1, b0 + k1, b0 + k))
    last = 1 + (rings - 2) * s
    for k in range(s):
        idx.append((bot, last + k, last + (k + 1) % s))
    pos = c + radius * dirs
    tris = orient_outward(pos, np.asarray(idx, dtype=np.int64), c)
    return Mesh(pos.astype(np.float32), dirs.astype(np.float32), tris.astype(np.int32))


def ground_quad(size_x: float, size_z: float, y: float = 0.0) -> Mesh:
    """Axis-aligned quad in the plane y, normal +y, two triangles."""
    hx, hz = size_x / 2.0, size_z / 2.0
    pos = np.array([[-hx, y, -hz], [hx, y, -hz], [hx, y, hz], [-hx, y, hz]], dtype=np.float32)
    nrm = np.tile(np.array([[0.0, 1.0, 0.0]], dtype=np.float32), (4, 1))
    idx = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.int32)  # CCW seen from +y
    return Mesh(pos, nrm, idx)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


def triangle_soup(n: int, seed: int, lo=(-5.0, 0.5, -5.0), hi=(5.0, 4.0, 5.0),
                  mu: float = -2.5, sigma: float = 0.5) -> Mesh:
    """`n` equilateral triangles: centroids uniform in [lo, hi], edge length
    lognormal(mu, sigma), uniformly random orientation; flat normals."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cen = rng.uniform(lo, hi, size=(n, 3))
    edge = rng.lognormal(mu, sigma, size=n)
    rot = _random_rotations(rng, n)
    r = edge / np.sqrt(3.0)  # circumradius of an equilateral triangle
    ang = np.array([0.0, 2.0 * np.pi / 3.0, 4.0 * np.pi / 3.0])
    local = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], axis=-1)  # (3,3)
    verts = cen[:, None, :] + np.einsum("nij,kj->nki", rot, local) * r[:, None, None]
    normal = rot[:, :, 2]
    pos = verts.reshape(-1, 3).astype(np.float32)
    nrm = np.repeat(normal, 3, axis=0).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return Mesh(pos, nrm, idx)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(freq: int):
    """Unit sphere from an icosahedron with each edge split `freq` times:
    20 * freq^2 triangles, shared vertices."""
    v, faces = _icosahedron()
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    keep = ii + jj <= freq
    ii, jj = ii[keep], jj[keep]
    lid = -np.ones((freq + 1, freq + 1), dtype=np.int64)
    lid[ii, jj] = np.arange(ii.size)
    tri_local = []
    for i in range(freq):
        for j in range(freq - i):
            tri_local.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j + 1 < freq:
                tri_local.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    tri_local = np.asarray(tri_local, dtype=np.int64)
    all_pts, all_tris = [], []
    for fi, (a, b, c) in enumerate(faces):
        pts = (v[a][None] * (freq - ii - jj)[:, None] + v[b][None] * ii[:, None] + v[c][None] * jj[:, None]) / freq
        all_pts.append(pts)
        all_tris.append(tri_local + fi * ii.size)
    pts = np.concatenate(all_pts)
    tris = np.concatenate(all_tris)
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    key = np.round(pts * 1e6).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    return pts[first], inverse[tris]


def value_noise_fbm(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0) -> np.ndarray:
    """Fractal (`octaves`, lacunarity 2, gain 1/2) of 3-D lattice value noise with
    smoothstep interpolation; lattice values uniform in [-1, 1] from `seed`.
    Normalised to [-1, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm])
    vals = rng.uniform(-1.0, 1.0, size=256)

    def lattice(ix, iy, iz):
        return vals[perm[perm[perm[ix & 255] + (iy & 255)] + (iz & 255)]]

    out = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        f = q - i0
        u = f * f * (3.0 - 2.0 * f)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            wx = u[:, 0] if dx else 1.0 - u[:, 0]
            for dy in (0, 1):
                wy = u[:, 1] if dy else 1.0 - u[:, 1]
                for dz in (0, 1):
                    wz = u[:, 2] if dz else 1.0 - u[:, 2]
                    acc += wx * wy * wz * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        out += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


def smooth_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals (Mesh::computeNormals,
    ShadowMapping/src/Mesh.cpp:195-234, accumulates face normals per vertex)."""
    p = pos.astype(np.float64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n


def displaced_icosphere(freq: int, seed: int, radius: float = 1.0, amplitude: float = 0.15,
                        center=(0.0, 1.2, 0.0)) -> Mesh:
    """Geodesic sphere (20*freq^2 triangles) displaced radially by
    `amplitude` * 4-octave value-noise fractal."""
    pts, tris = geodesic_sphere(freq)
    r = radius + amplitude * value_noise_fbm(pts, seed)
    pos = pts * r[:, None] + np.asarray(center)[None]
    tris = orient_outward(pos, tris, center)
    nrm = smooth_normals(pos, tris)
    return Mesh(pos.astype(np.float32), nrm.astype(np.float32), tris.astype(np.int32))


def stratified_jitter(n: int, seed: int) -> np.ndarray:
    """n = k*k stratified-jittered offsets in [-0.5, 0.5]^2 (row-major strata)."""
    k = int(round(n ** 0.5))
    if k * k != n:
        raise ValueError("stratified sampling needs a square sample count")
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.uniform(0.0, 1.0, size=(n, 2))
    ij = np.stack(np.meshgrid(np.arange(k), np.arange(k), indexing="xy"), -1).reshape(-1, 2)
    return ((ij + u) / k - 0.5).astype(np.float64)


# ---------------------------------------------------------------------------
# Light and camera descriptions
# ---------------------------------------------------------------------------
@dataclass
class View:
    """A pinhole view as MyGLGeometryViewer builds it: perspective(fov, W/H,
    1, 1000) * lookAt(eye, at, up) (MyGLGeometryViewer.cpp:6-26), or an
    orthographic box for directional lights."""
    eye: tuple
    at: tuple
    width: int
    height: int
    fov: float = 45.0
    znear: float = 1.0
    zfar: float = 1000.0
    up: tuple = UP
    ortho: tuple | None = None  # (l, r, b, t, n, f)


@dataclass
class Workload:
    name: str
    mesh: Mesh
    camera: View
    lights: list  # list[View], one per light sample
    technique: str
    shadow_intensity: float = 0.25
    zbias: float = 0.0
    extra: dict = field(default_factory=dict)


def ortho_fit(direction, mesh: Mesh, width: int, height: int, margin: float = 0.01) -> View:
    """Directional light: a view along `direction` whose orthographic box is
    the scene's bounds in light space (plus a relative margin)."""
    d = np.asarray(direction, dtype=np.float64)
    d /= np.linalg.norm(d)
    lo, hi = mesh.bounds()
    center = (lo.astype(np.float64) + hi) / 2.0
    radius = float(np.linalg.norm(hi.astype(np.float64) - lo)) / 2.0
    eye = center - d * (2.0 * radius + 1.0)
    from . import camera as cam
    v = cam.lookat(eye, center, UP).reshape(4, 4).T.astype(np.float64)
    corners = np.array([[x, y, z, 1.0] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    ls = corners @ v.T
    xmin, ymin = ls[:, 0].min(), ls[:, 1].min()
    xmax, ymax = ls[:, 0].max(), ls[:, 1].max()
    zn, zf = -ls[:, 2].max(), -ls[:, 2].min()
    mx, my, mz = margin * (xmax - xmin), margin * (ymax - ymin), margin * (zf - zn)
    box = (xmin - mx, xmax + mx, ymin - my, ymax + my, max(zn - mz, 1e-3), zf + mz)
    return View(tuple(eye), tuple(center), width, height, ortho=tuple(float(b) for b in box))


def config1(view_w: int = 800, view_h: int = 600, sm: int = 1024) -> Workload:
    """Hard shadow map: point light (4,8,4) -> origin, 1024^2 map, 800x600
    view, 3x3 PCF, constant bias 0.005; UV sphere r=1 (64x32) at y=1.2 over a
    20x20 quad at y=0; camera (0,3,6), fov 60."""
    mesh = Mesh.concat(uv_sphere(1.0, (0.0, 1.2, 0.0), 64, 32), ground_quad(20.0, 20.0, 0.0))
    light = View((4.0, 8.0, 4.0), (0.0, 0.0, 0.0), sm, sm)
    cam = View((0.0, 3.0, 6.0), (0.0, 0.0, 0.0), view_w, view_h, fov=60.0)
    return Workload("cfg1_hard_pcf", mesh, cam, [light], "pcf", zbias=0.005,
                    extra={"kernel_order": 3, "penumbra_size": 1})


def config2(ntris: int = 200_000, view_w: int = 1920, view_h: int = 1080, sm: int = 2048) -> Workload:
    """RBSM: 200k-triangle soup (seed 1) over a 40x40 quad, directional light
    along normalize(0.3,-1,0.2) with a fitted orthographic box."""
    soup = triangle_soup(ntris, seed=1)
    mesh = Mesh.concat(soup, ground_quad(40.0, 40.0, 0.0))
    light = ortho_fit((0.3, -1.0, 0.2), mesh, sm, sm)
    cam = View((0.0, 9.0, 14.0), (0.0, 1.0, 0.0), view_w, view_h, fov=60.0)
    return Workload("cfg2_rbsm", mesh, cam, [light], "rbsm",
                    extra={"depth_threshold": 2.5e-6, "max_search": 16, "kernel_order": 3, "penumbra_size": 1})


def area_light_samples(center, at, side: float, offsets: np.ndarray, sm_w: int, sm_h: int) -> list:
    """One View per sample: eye and look-at shifted by the same world-XY offset
    (UniformSampledLightSource.cpp:13-37 applies the offset to both)."""
    c = np.asarray(center, dtype=np.float64)
    a = np.asarray(at, dtype=np.float64)
    out = []
    for ox, oy in offsets:
        d = np.array([ox * side, oy * side, 0.0])
        out.append(View(tuple(c + d), tuple(a + d), sm_w, sm_h))
    return out


def config3(nsamples: int = 64, view_w: int = 1920, view_h: int = 1080, sm: int = 2048,
            freq: int = 71) -> Workload:
    """Area light (side 0.5, centre (3,7,2), 64 stratified-jittered samples,
    seed 2) over a ~100k-triangle value-noise icosphere (seed 3; 20*71^2 =
    100,820 triangles) above a 20x20 ground quad."""
    ico = displaced_icosphere(freq, seed=3)
    mesh = Mesh.concat(ico, ground_quad(20.0, 20.0, 0.0))
    offs = stratified_jitter(nsamples, seed=2)
    lights = area_light_samples((3.0, 7.0, 2.0), (0.0, 0.0, 0.0), 0.5, offs, sm, sm)
    cam = View((0.0, 3.0, 6.0), (0.0, 1.0, 0.0), view_w, view_h, fov=60.0)
    return Workload("cfg3_soft64", mesh, cam, lights, "soft", zbias=0.0,
                    extra={"offsets": offs, "side": 0.5, "kernel_order": 3, "penumbra": 1.0,
                           "center": (3.0, 7.0, 2.0), "at": (0.0, 0.0, 0.0)})


def config4(n: int = 32, seed: int = 4, map_sizes=(512, 1024, 2048, 4096),
            views=((640, 480), (1280, 720), (1920, 1080)), tri_range=(5000, 300000), max_samples: int = 16,
            soup_seed_base: int = 1000) -> list:
    """Batched frames: `n` independent instances (seed 4).  Each draws a square
    shadow-map size from `map_sizes`, a view from `views`, a triangle soup of
    log-uniform size in `tri_range` (over a 40x40 ground quad) and 1 to
    `max_samples` uniform samples of a 0.5-wide square light centred at a
    random point above the scene and looking at the origin."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        sm = int(map_sizes[int(rng.integers(len(map_sizes)))])
        vw, vh = views[int(rng.integers(len(views)))]
        lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
        ntri = int(round(float(np.exp(rng.uniform(lo, hi)))))
        ns = int(rng.integers(1, max_samples + 1))
        soup = triangle_soup(ntri, seed=soup_seed_base + i)
        mesh = Mesh.concat(soup, ground_quad(40.0, 40.0, 0.0))
        center = (float(rng.uniform(-4.0, 4.0)), float(rng.uniform(7.0, 10.0)), float(rng.uniform(-4.0, 4.0)))
        offs = rng.uniform(-0.5, 0.5, size=(ns, 2))
        lights = area_light_samples(center, (0.0, 0.0, 0.0), 0.5, offs, sm, sm)
        cam = View((0.0, 9.0, 14.0), (0.0, 1.0, 0.0), int(vw), int(vh), fov=60.0)
        out.append(Workload(f"cfg4_{i}", mesh, cam, lights, "soft", extra={"triangles": ntri, "samples": ns}))
    return out


def config5(nsamples: int = 16, view: int = 512, sm: int = 1024, freq: int = 50) -> Workload:
    """Differentiable soft shadows: 16-sample light, 1024^2 maps, 512^2 view,
    50k-triangle displaced icosphere (seed 5; 20*50^2 = 50,000 triangles)."""
    ico = displaced_icosphere(freq, seed=5)
    mesh = Mesh.concat(ico, ground_quad(20.0, 20.0, 0.0))
    offs = stratified_jitter(nsamples, seed=5)
    lights = area_light_samples((3.0, 7.0, 2.0), (0.0, 0.0, 0.0), 0.5, offs, sm, sm)
    cam = View((0.0, 3.0, 6.0), (0.0, 1.0, 0.0), view, view, fov=60.0)
    return Workload("cfg5_diff", mesh, cam, lights, "soft",
                    extra={"offsets": offs, "side": 0.5, "center": (3.0, 7.0, 2.0), "at": (0.0, 0.0, 0.0)})


def _load(handle, name: str, window=(1024, 1024), sm: int = 2048) -> Workload:
    import ctypes

    from . import _lib

    L = _lib.lib()
    try:
        info = _lib.GiSceneInfo()
        _lib.call("gi_scene_get_info", handle, ctypes.byref(info))
        nv, nt = int(info.nverts), int(info.ntris)
        pos = np.empty((nv, 3), np.float32)
        nrm = np.empty((nv, 3), np.float32)
        idx = np.empty((nt, 3), np.int32)
        col = np.empty((nv, 3), np.float32)
        _lib.call("gi_scene_copy", handle, pos.ctypes.data, nrm.ctypes.data, idx.ctypes.data, col.ctypes.data)
        textures = [L.gi_scene_texture_path(handle, i).decode() for i in range(int(info.ntextures))]
    finally:
        L.gi_scene_free(handle)
    cam = View(tuple(info.camera_eye), tuple(info.camera_at), int(window[0]), int(window[1]), fov=45.0,
               up=(0.0, 0.0, 1.0))
    light = View(tuple(info.light_eye), tuple(info.light_at), int(sm), int(sm), fov=45.0, up=(0.0, 0.0, 1.0))
    extra = {"depth_threshold": float(info.depth_threshold), "hsm": (float(info.hsm_alpha), float(info.hsm_beta)),
             "colors": col if info.has_colors else None, "textures": textures}
    return Workload(name, Mesh(pos, nrm, idx), cam, [light], "hard", extra=extra)


def load_scene(config_path, base_dir=None, window=(1024, 1024), sm: int = 2048) -> Workload:
    """A reference scene configuration (Configs/<name>.txt, format of
    ShadowMapping/src/IO/SceneLoader.cpp:12-91) read by libgi's loader:
    the merged mesh (OBJ files with recomputed normals, scaled, translated
    and rotated in file order), the camera (`ve`/`va`) and light (`le`/`la`)
    with up (0, 0, 1) and fov 45 (ShadowMapping/src/main.cpp:863-867), the
    depth threshold and per-vertex colours.  Relative paths resolve against
    `base_dir` (the reference's working directory is its repository root)."""
    import ctypes
    from pathlib import Path

    from . import _lib

    h = ctypes.c_void_p()
    base = None if base_dir is None else str(base_dir).encode()
    _lib.call("gi_scene_load", str(config_path).encode(), base, ctypes.byref(h))
    return _load(h, Path(str(config_path)).stem, window, sm)


def load_obj(obj_path) -> Mesh:
    """One OBJ file as the loader's `o` line reads it (Mesh::loadOBJFile +
    computeNormals, ShadowMapping/src/Mesh.cpp:195-313)."""
    import ctypes

    from . import _lib

    h = ctypes.c_void_p()
    _lib.call("gi_obj_load", str(obj_path).encode(), ctypes.byref(h))
    return _load(h, "obj").mesh

