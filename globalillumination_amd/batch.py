"""Batched shadow frames over the gi_batch_* C ABI.

An instance is the work of one renderMonteCarlo frame
(SoftShadowMapping/src/main.cpp:756-811): the depth maps of its light samples
and the soft-shadow screen pass over its camera G-buffer.  A Batch runs many
instances (different meshes, map sizes, sample counts and views) as one
sequence of persistent kernels, optionally captured in one hipGraph; every
map and shadow value equals the per-instance passes (render_shadow_map +
render_monte_carlo) bit for bit.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .passes import DeviceMesh, GBuffer, SoftParams, _dptr, _stream_handle, render_monte_carlo, render_shadow_map, \
    soft_uniforms

SHAPES = {0: (64, 4), 1: (32, 8), 2: (16, 16), 3: "16x16 of 8x8 wave tiles", 4: "64x4 of 16x4 wave tiles"}


@dataclass
class BatchInstance:
    mesh: DeviceMesh
    light_mvps: np.ndarray  # (S, 16) unbiased P*V per sample, column-major FP32
    map_width: int
    map_height: int
    gbuffer: GBuffer
    soft: SoftParams
    zbias: float = 0.0
    polygon_offset: tuple = (0.0, 0.0)
    maps: torch.Tensor = field(default=None)  # (S, H, W) output
    out: torch.Tensor = field(default=None)   # (view H, view W) output

    def __post_init__(self):
        dev = self.mesh.positions.device
        self.light_mvps = np.ascontiguousarray(self.light_mvps, dtype=np.float32).reshape(-1, 16)
        S = self.light_mvps.shape[0]
        if self.maps is None:
            self.maps = torch.empty((S, self.map_height, self.map_width), dtype=torch.float32, device=dev)
        if self.out is None:
            self.out = torch.empty((self.gbuffer.height, self.gbuffer.width), dtype=torch.float32, device=dev)
        self.common, trans = soft_uniforms(self.light_mvps, self.zbias)
        self.mvps_dev = torch.from_numpy(self.light_mvps).to(dev)
        self.trans_dev = torch.from_numpy(np.ascontiguousarray(trans)).to(dev)

    @property
    def nsamples(self) -> int:
        return int(self.light_mvps.shape[0])

    def to_c(self) -> _lib.GiBatchInstance:
        c = _lib.GiBatchInstance()
        c.positions_dev = _dptr(self.mesh.positions).value
        c.nverts = self.mesh.nverts
        c.indices_dev = _dptr(self.mesh.indices).value if self.mesh.ntris else None
        c.ntris = self.mesh.ntris
        c.light_mvps_dev = _dptr(self.mvps_dev).value
        c.nsamples = self.nsamples
        c.sm_width, c.sm_height = int(self.map_width), int(self.map_height)
        c.polygon_offset_factor, c.polygon_offset_units = (float(x) for x in self.polygon_offset)
        c.maps_dev = _dptr(self.maps).value
        c.gbuffer = self.gbuffer.to_c(with_normals=False)
        c.light_mvp = _lib.F16(*np.asarray(self.common, dtype=np.float32).reshape(-1).tolist())
        c.trans_dev = _dptr(self.trans_dev).value
        c.soft = self.soft.to_c()
        c.out_dev = _dptr(self.out).value
        return c

    def run_unbatched(self, stream=None) -> None:
        """The same frame through the single-instance passes."""
        render_shadow_map(self.mesh, self.light_mvps, self.map_width, self.map_height,
                          polygon_offset=self.polygon_offset, out=self.maps, stream=stream)
        render_monte_carlo(self.gbuffer, self.maps, self.light_mvps, self.soft, zbias=self.zbias, out=self.out,
                           trans_dev=self.trans_dev, stream=stream)


class BatchGraph:
    """A Batch's launches captured in a hipGraph (the tuning at capture)."""

    def __init__(self, batch: "Batch"):
        h = ctypes.c_void_p()
        _lib.call("gi_batch_graph_create", batch.handle, ctypes.byref(h))
        self.handle = h
        self.batch = batch  # keeps the plan and its buffers alive

    def launch(self, stream=None) -> None:
        _lib.call("gi_batch_graph_launch", self.handle, _stream_handle(stream))

    def close(self) -> None:
        if self.handle:
            _lib.lib().gi_batch_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Plan over a list of BatchInstance (device tables in its own workspace)."""

    def __init__(self, instances: list):
        if not instances:
            raise ValueError("a batch needs at least one instance")
        self.instances = list(instances)
        n = len(self.instances)
        self._c = (_lib.GiBatchInstance * n)(*[i.to_c() for i in self.instances])
        nbytes = _lib.lib().gi_batch_workspace_size(self._c, n)
        if nbytes == 0:
            raise _lib.GiError("gi_batch_workspace_size", _lib.GI_EINVAL, "invalid instance")
        dev = self.instances[0].mesh.positions.device
        self.workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        h = ctypes.c_void_p()
        _lib.call("gi_batch_create", self._c, n, _dptr(self.workspace), ctypes.c_size_t(int(nbytes)),
                  ctypes.byref(h))
        self.handle = h

    @property
    def tuning(self) -> dict:
        t = _lib.GiBatchTuning()
        _lib.call("gi_batch_get_tuning", self.handle, ctypes.byref(t))
        return {"tile": t.tile, "raster_groups": t.raster_groups, "screen_shape": SHAPES[t.screen_shape],
                "screen_groups": t.screen_groups}

    def set_tuning(self, tile=64, raster_groups=4, screen_shape=3, screen_groups=4) -> None:
        t = _lib.GiBatchTuning(int(tile), int(raster_groups), int(screen_shape), int(screen_groups))
        _lib.call("gi_batch_set_tuning", self.handle, ctypes.byref(t))

    def run(self, stream=None) -> None:
        _lib.call("gi_batch_run", self.handle, _stream_handle(stream))

    def autotune(self, reps: int = 3, stream=None) -> tuple:
        """Measures every candidate and keeps the fastest: (tuning, ms)."""
        t = _lib.GiBatchTuning()
        ms = ctypes.c_float()
        _lib.call("gi_batch_autotune", self.handle, _stream_handle(stream), int(reps), ctypes.byref(t),
                  ctypes.byref(ms))
        return self.tuning, float(ms.value)

    def graph(self) -> BatchGraph:
        return BatchGraph(self)

    def close(self) -> None:
        if self.handle:
            _lib.lib().gi_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
