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
