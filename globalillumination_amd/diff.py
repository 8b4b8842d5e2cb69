"""Differentiable area-light soft shadows: a torch.autograd.Function over the
adjoint kernels of libgi (gi_diff_soft_forward / gi_diff_soft_backward).

Parameters: the light position L (3), the light radius r (1) and every mesh
vertex.  Light sample i sits at eye_i = L + r*(ox_i, oy_i, 0) and looks at
A + r*(ox_i, oy_i, 0): eye and look-at shift together, as
UniformSampledLightSource places its point lights
(SoftShadowMapping/src/Scene/LightSource/UniformSampledLightSource.cpp:13-37),
with offsets o_i in [-1, 1]^2 (the light square has half-side r).  Each
sample's matrix is P * lookAt(eye_i, at_i, up) (MyGLGeometryViewer.cpp:14-26).

The kernels return dL/d(matrices) and dL/d(vertices); the chain from the
matrices to (L, r) is taken here in float64 with complex-step derivatives of
glm's lookAt/perspective.  The camera G-buffer is a fixed input, as in
renderMonteCarlo, which reuses the last G-buffer (SoftShadowMapping/src/main.cpp:756-811).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .passes import GBuffer, Workspace, _dptr, _stream_handle

BIAS = np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0, 0.5], [0, 0, 0.5, 0.5], [0, 0, 0, 1.0]])  # B, row-major


# ---------------------------------------------------------------------------
# glm formulas in float64 / complex128 (row-major math matrices)
# ---------------------------------------------------------------------------
def _normalize(v):
    return v / np.sqrt(np.sum(v * v))


def lookat64(eye, at, up):
    eye, at, up = (np.asarray(x) for x in (eye, at, up))
    f = _normalize(at - eye)
    u = _normalize(up.astype(f.dtype))
    s = _normalize(np.cross(f, u))
    u = np.cross(s, f)
    m = np.zeros((4, 4), dtype=f.dtype)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[3, 3] = 1.0
    m[0, 3], m[1, 3], m[2, 3] = -np.dot(s, eye), -np.dot(u, eye), np.dot(f, eye)
    return m


def perspective64(fovy_deg, aspect, zn, zf):
    t = np.tan(np.radians(fovy_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = -(zf + zn) / (zf - zn)
    m[2, 3] = -2.0 * zf * zn / (zf - zn)
    m[3, 2] = -1.0
    return m


def colmajor(m):
    return np.ascontiguousarray(np.asarray(m).T.reshape(-1))


@dataclass
class AreaLight:
    at: tuple             # look-at centre A
    offsets: np.ndarray   # (N, 2) in [-1, 1]
    map_size: int
    fov: float = 45.0
    znear: float = 1.0
    zfar: float = 1000.0
    up: tuple = (0.0, 0.0, 1.0)

    @staticmethod
    def from_workload(wl) -> tuple:
        """(AreaLight, L, r) of a workload built by area_light_samples: the
        square of side `side` has half-side r and unit offsets 2*jitter."""
        ex = wl.extra
        light = AreaLight(tuple(ex["at"]), 2.0 * np.asarray(ex["offsets"], np.float64), wl.lights[0].width,
                          fov=wl.lights[0].fov, znear=wl.lights[0].znear, zfar=wl.lights[0].zfar,
                          up=tuple(wl.lights[0].up))
        return light, np.asarray(ex["center"], np.float64), 0.5 * float(ex["side"])

    def matrices(self, L, r):
        """Unbiased P*V_i of every sample, row-major, dtype of L/r (float64 or
        complex128 for complex-step differentiation)."""
        L = np.asarray(L)
        P = perspective64(self.fov, 1.0, self.znear, self.zfar)
        out = []
        for ox, oy in self.offsets:
            d = r * np.array([ox, oy, 0.0])
            out.append(P @ lookat64(L + d, np.asarray(self.at) + d, np.asarray(self.up)))
        return np.stack(out)

    def jacobian(self, L, r, h=1e-30):
        """d(P*V_i)/d(Lx, Ly, Lz, r), shape (N, 4, 4, 4): complex-step, exact
        to float64 rounding."""
        th = np.array([*L, r], dtype=np.complex128)
        J = []
        for k in range(4):
            t = th.copy()
            t[k] += 1j * h
            J.append(self.matrices(t[:3], t[3]).imag / h)
        return np.stack(J, axis=-1)


@dataclass
class DiffParams:
    tau: float = 2e-4
    bias: float = 2e-3
    shadow_intensity: float = 0.25

    def to_c(self):
        return _lib.GiDiffParams(float(self.tau), float(self.bias), float(self.shadow_intensity))


_WS = Workspace()


class DiffScene:
    """Fixed inputs of the differentiable pass: the mesh topology, the camera
    G-buffer, the area-light description and the model parameters."""

    def __init__(self, indices: torch.Tensor, gbuffer: GBuffer, light: AreaLight, params: DiffParams = DiffParams()):
        self.indices = indices.contiguous()
        self.gbuffer = gbuffer
        self.light = light
        self.params = params
        self._mcache = None

    @property
    def S(self):
        return len(self.light.offsets)

    def device_matrices(self, L, r, device):
        """FP64 column-major (unbiased, biased) matrices of every sample on
        `device`; the last pair is cached."""
        key = (tuple(np.asarray(L, np.float64).tolist()), float(r), str(device))
        if self._mcache is not None and self._mcache[0] == key:
            return self._mcache[1], self._mcache[2]
        M = self.light.matrices(np.asarray(L, np.float64), float(r))
        Mb = np.einsum("ij,njk->nik", BIAS, M)
        mv = torch.from_numpy(np.stack([colmajor(m) for m in M])).to(device)
        bv = torch.from_numpy(np.stack([colmajor(m) for m in Mb])).to(device)
        self._mcache = (key, mv, bv)
        return mv, bv

    def _workspace(self, device):
        sm = self.light.map_size
        n = _lib.lib().gi_diff_workspace_size(int(self.indices.shape[0]), self.S, sm, sm, self.gbuffer.width,
                                              self.gbuffer.height)
        if n == 0:
            raise _lib.GiError("gi_diff_workspace_size", _lib.GI_EINVAL, "unsupported sizes")
        return _WS.get(n, device)

    def forward(self, L, r, vertices: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        dev = vertices.device
        mv, bv = self.device_matrices(L, r, dev)
        v = vertices.detach().contiguous()
        if out is None:
            out = torch.empty((self.gbuffer.height, self.gbuffer.width), dtype=torch.float32, device=dev)
        ws = self._workspace(dev)
        g = self.gbuffer.to_c(with_normals=False)
        sm = self.light.map_size
        _lib.call("gi_diff_soft_forward", _dptr(v), int(v.shape[0]), _dptr(self.indices), int(self.indices.shape[0]),
                  _dptr(mv), _dptr(bv), self.S, sm, sm, ctypes.byref(g), ctypes.byref(self.params.to_c()),
                  _dptr(out), _dptr(ws), ctypes.c_size_t(ws.numel()), _stream_handle())
        return out

    def backward_device(self, L, r, vertices: torch.Tensor, grad_out: torch.Tensor):
        """Launches the adjoint; returns device tensors (dL/d(B*M'_i),
        dL/dM'_i as (S, 16) column-major, dL/dvertices (nv, 3))."""
        dev = vertices.device
        mv, bv = self.device_matrices(L, r, dev)
        v = vertices.detach().contiguous()
        go = grad_out.detach().contiguous().float()
        gb_m = torch.empty((self.S, 16), dtype=torch.float64, device=dev)
        g_m = torch.empty((self.S, 16), dtype=torch.float64, device=dev)
        gv = torch.zeros_like(v)
        ws = self._workspace(dev)
        g = self.gbuffer.to_c(with_normals=False)
        sm = self.light.map_size
        _lib.call("gi_diff_soft_backward", _dptr(v), int(v.shape[0]), _dptr(self.indices),
                  int(self.indices.shape[0]), _dptr(mv), _dptr(bv), self.S, sm, sm, ctypes.byref(g),
                  ctypes.byref(self.params.to_c()), _dptr(go), _dptr(gb_m), _dptr(g_m), _dptr(gv), _dptr(ws),
                  ctypes.c_size_t(ws.numel()), _stream_handle())
        return gb_m, g_m, gv

    def light_gradient(self, L, r, gb_m: torch.Tensor, g_m: torch.Tensor) -> np.ndarray:
        """dL/d(Lx, Ly, Lz, r) in float64 from the matrix adjoints:
        dL/dM'_i = dL/dM'_i (depth maps) + B^T dL/d(B M'_i) (receivers)."""
        Gb = gb_m.cpu().numpy().reshape(self.S, 4, 4).transpose(0, 2, 1)
        Gm = g_m.cpu().numpy().reshape(self.S, 4, 4).transpose(0, 2, 1)
        G = Gm + np.einsum("ji,njk->nik", BIAS, Gb)
        J = self.light.jacobian(np.asarray(L, np.float64), float(r))
        return np.einsum("nij,nijk->k", G, J)

    def backward(self, L, r, vertices: torch.Tensor, grad_out: torch.Tensor):
        """(dL/dL (3,), dL/dr, dL/dvertices (nv, 3))."""
        gb_m, g_m, gv = self.backward_device(L, r, vertices, grad_out)
        gtheta = self.light_gradient(L, r, gb_m, g_m)
        return gtheta[:3], gtheta[3], gv


class SoftShadowFunction(torch.autograd.Function):
    """S = soft_shadow(light_position, light_radius, vertices) for a fixed
    DiffScene; backward recomputes the depth pass and runs the adjoint."""

    @staticmethod
    def forward(ctx, light_position: torch.Tensor, light_radius: torch.Tensor, vertices: torch.Tensor,
                scene: DiffScene):
        L = light_position.detach().double().cpu().numpy()
        r = float(light_radius.detach().double().cpu())
        ctx.scene, ctx.L, ctx.r = scene, L, r
        ctx.save_for_backward(vertices)
        return scene.forward(L, r, vertices)

    @staticmethod
    def backward(ctx, grad_out):
        (vertices,) = ctx.saved_tensors
        gL, gr, gv = ctx.scene.backward(ctx.L, ctx.r, vertices, grad_out)
        dev = grad_out.device
        return (torch.tensor(gL, dtype=torch.float32, device=dev),
                torch.tensor(gr, dtype=torch.float32, device=dev).reshape(1), gv, None)


def soft_shadow(light_position, light_radius, vertices, scene: DiffScene):
    return SoftShadowFunction.apply(light_position, light_radius, vertices, scene)
