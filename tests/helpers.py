"""Shared helpers of the parity tests: build workloads, run the HIP path and
the oracle on identical inputs."""
from __future__ import annotations

import numpy as np


def light_mvps(workload) -> np.ndarray:
    from globalillumination_amd import passes

    return np.stack([passes.view_matrices(v)[2] for v in workload.lights])


def camera_mvp(workload) -> np.ndarray:
    from globalillumination_amd import passes

    return passes.view_matrices(workload.camera)[2]


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and bool((a.view(np.uint32) == b.view(np.uint32)).all())


def mismatch_report(a: np.ndarray, b: np.ndarray, name: str = "") -> str:
    d = a.view(np.uint32) != b.view(np.uint32)
    idx = np.argwhere(d)
    n = int(d.sum())
    head = idx[:5].tolist()
    vals = [(float(a[tuple(i)]), float(b[tuple(i)])) for i in idx[:5]]
    return f"{name}: {n} of {d.size} differ; first {head} values {vals}"


def shrink_soup(n: int, seed: int = 1):
    from globalillumination_amd import scenes

    soup = scenes.triangle_soup(n, seed=seed)
    return scenes.Mesh.concat(soup, scenes.ground_quad(40.0, 40.0, 0.0))


def clip_heavy_mesh(n_small: int = 8000, n_clip: int = 200, seed: int = 3):
    """n_small tiny triangles under a light at (0, 10, 0.1) and n_clip large
    ones around the light's eye, most of which straddle the near plane or the
    guard band and are cut into fan pieces by the clip kernel."""
    from globalillumination_amd import scenes

    rng = np.random.default_rng(seed)
    eye = np.array([0.0, 10.0, 0.1], np.float32)
    big = eye + rng.uniform(-15, 15, size=(n_clip, 3, 3))
    cen = np.c_[rng.uniform(-3, 3, n_small), rng.uniform(0, 2, n_small), rng.uniform(-3, 3, n_small)]
    small = cen[:, None, :] + rng.uniform(-0.05, 0.05, size=(n_small, 3, 3))
    pos = np.concatenate([small, big]).astype(np.float32)
    n = n_small + n_clip
    nrm = np.tile(np.array([0, 1, 0], np.float32), (3 * n, 1))
    mesh = scenes.Mesh(pos.reshape(-1, 3), nrm, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    light = scenes.View((0.0, 10.0, 0.1), (0.0, 0.0, 0.0), 128, 128, fov=60.0)
    return mesh, light


def overflow_mesh(n: int = 64, seed: int = 7, size: int = 1024):
    """n large overlapping triangles under a light at (0, 50, 0.1), fov 90:
    every one touches every tile, so the tile lists exceed the workspace
    bound and the map goes through the fallback raster."""
    from globalillumination_amd import scenes

    rng = np.random.default_rng(seed)
    pos = rng.uniform(-30, 30, size=(n, 3, 3)).astype(np.float32)
    pos[:, :, 1] = rng.uniform(0.0, 2.0, size=(n, 3))
    nrm = np.tile(np.array([0, 1, 0], np.float32), (3 * n, 1))
    mesh = scenes.Mesh(pos.reshape(-1, 3), nrm, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    light = scenes.View((0.0, 50.0, 0.1), (0.0, 0.0, 0.0), size, size, fov=90.0)
    return mesh, light


def wide_overflow_mesh(n: int = 3000, seed: int = 11, size: int = 512):
    """n mid-sized triangles under a light at (0, 40, 0.1), each wider than
    the binning's per-thread span: more setup records than the workspace
    holds, so the map goes through the fallback raster."""
    from globalillumination_amd import scenes

    rng = np.random.default_rng(seed)
    cen = rng.uniform(-6, 6, size=(n, 1, 3))
    pos = (cen + rng.uniform(-3, 3, size=(n, 3, 3))).astype(np.float32)
    pos[:, :, 1] = rng.uniform(0.0, 3.0, size=(n, 3))
    nrm = np.tile(np.array([0, 1, 0], np.float32), (3 * n, 1))
    mesh = scenes.Mesh(pos.reshape(-1, 3), nrm, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    light = scenes.View((0.0, 40.0, 0.1), (0.0, 0.0, 0.0), size, size, fov=45.0)
    return mesh, light


def clipped_count(mesh, mvp, guard=4.0):
    """Triangles the depth pass lists for clipping: not outside any of the six
    planes (z + w, w - z, x + g w, g w - x, y + g w, g w - y >= 0, g the guard
    band) and not inside all of them."""
    M = np.asarray(mvp, np.float64).reshape(4, 4).T
    h = np.c_[mesh.positions.astype(np.float64), np.ones(mesh.nverts)] @ M.T
    x, y, z, w = h.T
    d = np.stack([z + w, w - z, x + guard * w, guard * w - x, y + guard * w, guard * w - y], 1)[mesh.indices]
    reject = (d < 0).all(axis=1).any(axis=1)
    inside = (d >= 0).all(axis=(1, 2))
    return int((~reject & ~inside).sum())
