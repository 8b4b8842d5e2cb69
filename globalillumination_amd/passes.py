"""Host-side pass interface over libgi.so, mirroring the reference's passes.

The reference drives each technique with free pass functions in its apps'
main.cpp and a ``ShadowParams`` struct handed to MyGLGeometryViewer::configure*:

    renderShadowMap      ShadowMapping/src/main.cpp:350-361
    renderGBuffer        ShadowMapping/src/main.cpp:363-372
    computeHardShadows   ShadowMapping/src/main.cpp:400-414
    renderMonteCarlo     SoftShadowMapping/src/main.cpp:756-811

The functions below keep those names (in snake case) and that order of work.
Device memory and streams come from torch; every pass runs a HIP kernel of
libgi.so on the current torch stream, and there is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib, camera
from .scenes import Mesh, View

TECHNIQUES = {"naive": _lib.GI_NAIVE, "pcf": _lib.GI_PCF, "rbsm": _lib.GI_RBSM,
              "rpcf": _lib.GI_RPCF, "rbsm_filtered": _lib.GI_RBSM_FILTERED,
              "rbsm_conservative": _lib.GI_RBSM_CONSERVATIVE, "rpcf_conservative": _lib.GI_RPCF_CONSERVATIVE}


def _stream_handle(stream=None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dptr(t: torch.Tensor) -> ctypes.c_void_p:
    """Device pointer of a contiguous tensor (NULL for an empty one)."""
    if not t.is_cuda:
        raise ValueError("libgi passes take device tensors")
    if not t.is_contiguous():
        raise ValueError("libgi passes take contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def _f(arr, n) -> ctypes.Array:
    a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
    if a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return (ctypes.c_float * n)(*a.tolist())


class Workspace:
    """Grow-only device scratch owned by the caller side (the library never
    allocates)."""

    def __init__(self):
        self._buf: dict = {}

    def get(self, nbytes: int, device) -> torch.Tensor:
        key = str(device)
        buf = self._buf.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            self._buf[key] = buf
        return buf


_WS = Workspace()


@dataclass
class DeviceMesh:
    positions: torch.Tensor  # (nv, 3) float32
    normals: torch.Tensor    # (nv, 3) float32
    indices: torch.Tensor    # (nt, 3) int32

    @staticmethod
    def from_mesh(m: Mesh, device="cuda") -> "DeviceMesh":
        return DeviceMesh(torch.from_numpy(np.ascontiguousarray(m.positions)).to(device),
                          torch.from_numpy(np.ascontiguousarray(m.normals)).to(device),
                          torch.from_numpy(np.ascontiguousarray(m.indices)).to(device))

    @property
    def nverts(self) -> int:
        return int(self.positions.shape[0])

    @property
    def ntris(self) -> int:
        return int(self.indices.shape[0])


# ---------------------------------------------------------------------------
# Matrices (MyGLGeometryViewer::configureAmbient / configureShadow)
# ---------------------------------------------------------------------------
def view_matrices(v: View):
    """(P, V, P*V) of a view; model = identity (MyGLGeometryViewer.cpp:14-26)."""
    if v.ortho is not None:
        P = camera.ortho(*v.ortho)
    else:
        P = camera.perspective(v.fov, v.width / v.height, v.znear, v.zfar)
    V = camera.lookat(v.eye, v.at, v.up)
    return P, V, camera.mat4_mul(P, V)


def biased(mvp: np.ndarray, zbias: float = 0.0) -> np.ndarray:
    """B * lightMVP (MyGLGeometryViewer.cpp:139-146), constant bias in B[3][2]."""
    return camera.mat4_mul(camera.bias_matrix(zbias), mvp)


@dataclass
class ShadowParams:
    """Mirror of ShadowParams (ShadowMapping/include/Viewers/ShadowParams.h:6-37)
    restricted to the fields the hard-shadow passes read."""
    lightMVP: np.ndarray                 # unbiased P*V*M of the light
    shadowMapWidth: int
    shadowMapHeight: int
    technique: str = "naive"
    kernelOrder: int = 3
    penumbraSize: int = 1
    maxSearch: int = 16
    depthThreshold: float = 2.5e-6
    shadowIntensity: float = 0.25
    zbias: float = 0.0
    preEvaluation: bool = False
    MV: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32).reshape(-1))
    normalMatrix: np.ndarray = field(default_factory=lambda: np.eye(3, dtype=np.float32).reshape(-1))
    lightPosition: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> _lib.GiShadowParams:
        p = _lib.GiShadowParams()
        p.light_mvp = _f(biased(self.lightMVP, self.zbias), 16)
        p.shadow_map_width = int(self.shadowMapWidth)
        p.shadow_map_height = int(self.shadowMapHeight)
        p.technique = TECHNIQUES[self.technique]
        p.kernel_order = int(self.kernelOrder)
        p.penumbra_size = int(self.penumbraSize)
        p.max_search = int(self.maxSearch)
        p.depth_threshold = float(self.depthThreshold)
        p.shadow_intensity = float(self.shadowIntensity)
        p.pre_evaluation = 1 if self.preEvaluation else 0
        p.mv = _f(self.MV, 16)
        p.normal_matrix = _f(self.normalMatrix, 9)
        p.light_position = _f(self.lightPosition, 3)
        return p


def phong_uniforms(cam: View, light_eye) -> tuple:
    """MV, normalMatrix and lightPosition as displaySceneFromGBuffer uploads
    them: lightPosition is the world light rotated 180 deg about y
    (ShadowMapping/src/main.cpp:309, configurePhong MyGLGeometryViewer.cpp:108-134)."""
    MV = camera.lookat(cam.eye, cam.at, cam.up)
    nm = camera.normal_matrix(MV)
    lx, ly, lz = (float(c) for c in light_eye)
    return MV, nm, (-lx, ly, -lz)


# ---------------------------------------------------------------------------
# Passes
# ---------------------------------------------------------------------------
def render_shadow_map(mesh: DeviceMesh, mvps: np.ndarray, width: int, height: int,
                      polygon_offset=(0.0, 0.0), out: torch.Tensor | None = None,
                      stream=None, raster_events=None) -> torch.Tensor:
    """Depth of `mesh` from each light in `mvps` ((S, 16) unbiased P*V*M) into
    an (S, height, width) FP32 tensor (renderShadowMap, and renderMonteCarlo's
    per-sample loop SoftShadowMapping/src/main.cpp:774-793).  `raster_events`:
    optional pair of torch.cuda.Event (enable_timing) recorded around the
    tile-raster kernel.  `mvps` may also be a device tensor of (S, 16) FP32
    matrices already resident on the mesh's device: then nothing is copied
    and the host does not wait on the stream."""
    dev = mesh.positions.device
    if isinstance(mvps, torch.Tensor):
        if mvps.device != dev or mvps.dtype != torch.float32 or not mvps.is_contiguous() or mvps.numel() % 16:
            raise ValueError("device mvps must be a contiguous float32 (S, 16) tensor on the mesh's device")
        mvps_dev = mvps
        S = mvps.numel() // 16
    else:
        m = np.ascontiguousarray(mvps, dtype=np.float32).reshape(-1, 16)
        S = m.shape[0]
        mvps_dev = torch.from_numpy(m).to(dev)
    if out is None:
        out = torch.empty((S, height, width), dtype=torch.float32, device=dev)
    nbytes = _lib.lib().gi_depth_pass_workspace_size(mesh.ntris, S, width, height)
    if nbytes == 0:
        raise _lib.GiError("gi_depth_pass_workspace_size", _lib.GI_EINVAL, "bad size")
    ws = _WS.get(nbytes, dev)
    ev = (None, None) if raster_events is None else (raster_events[0].cuda_event, raster_events[1].cuda_event)
    opts = _lib.GiRasterOpts(float(polygon_offset[0]), float(polygon_offset[1]), ev[0], ev[1])
    _lib.call("gi_depth_pass", _dptr(mesh.positions), mesh.nverts, _dptr(mesh.indices), mesh.ntris,
              _dptr(mvps_dev), S, width, height, ctypes.byref(opts), _dptr(out), _dptr(ws),
              ctypes.c_size_t(ws.numel()), _stream_handle(stream))
    return out


class SplitDepthPass:
    """gi_depth_pass in two steps (gi_depth_pass_bin, then
    gi_depth_pass_raster over ranges of maps), with a workspace of its own so
    that the steps may interleave with other passes on other streams.
    `mvps`: (S, 16) unbiased matrices, host array or device tensor."""

    def __init__(self, mesh: DeviceMesh, mvps, width: int, height: int, polygon_offset=(0.0, 0.0),
                 out: torch.Tensor | None = None):
        dev = mesh.positions.device
        if isinstance(mvps, torch.Tensor):
            if mvps.device != dev or mvps.dtype != torch.float32 or not mvps.is_contiguous() or mvps.numel() % 16:
                raise ValueError("device mvps must be a contiguous float32 (S, 16) tensor on the mesh's device")
            self.mvps_dev = mvps
        else:
            self.mvps_dev = torch.from_numpy(np.ascontiguousarray(mvps, dtype=np.float32).reshape(-1, 16)).to(dev)
        self.mesh, self.S, self.width, self.height = mesh, self.mvps_dev.numel() // 16, int(width), int(height)
        self.polygon_offset = (float(polygon_offset[0]), float(polygon_offset[1]))
        self.out = out if out is not None else torch.empty((self.S, height, width), dtype=torch.float32, device=dev)
        nbytes = _lib.lib().gi_depth_pass_workspace_size(mesh.ntris, self.S, self.width, self.height)
        if nbytes == 0:
            raise _lib.GiError("gi_depth_pass_workspace_size", _lib.GI_EINVAL, "bad size")
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)

    def _args(self, events=None):
        ev = (None, None) if events is None else (events[0].cuda_event, events[1].cuda_event)
        opts = _lib.GiRasterOpts(self.polygon_offset[0], self.polygon_offset[1], ev[0], ev[1])
        m = self.mesh
        return (_dptr(m.positions), m.nverts, _dptr(m.indices), m.ntris, _dptr(self.mvps_dev), self.S, self.width,
                self.height, ctypes.byref(opts))

    def bin(self, stream=None) -> None:
        _lib.call("gi_depth_pass_bin", *self._args(), _dptr(self.out), _dptr(self.ws), ctypes.c_size_t(self.ws.numel()),
                  _stream_handle(stream))

    def raster(self, map_begin: int, map_count: int, stream=None, raster_events=None) -> None:
        _lib.call("gi_depth_pass_raster", *self._args(raster_events), int(map_begin), int(map_count), _dptr(self.out),
                  _dptr(self.ws), ctypes.c_size_t(self.ws.numel()), _stream_handle(stream))


@dataclass
class GBuffer:
    position: torch.Tensor  # (3, H, W)
    normal: torch.Tensor    # (3, H, W), facing-flipped
    facing: torch.Tensor | None = None  # (H, W): 1.0 front-facing, 0.0 back-facing or background

    @property
    def width(self) -> int:
        return int(self.position.shape[2])

    @property
    def height(self) -> int:
        return int(self.position.shape[1])

    def to_c(self, with_normals: bool = True) -> _lib.GiGBuffer:
        return _lib.GiGBuffer(_dptr(self.position).value,
                              _dptr(self.normal).value if with_normals else None,
                              self.width, self.height,
                              _dptr(self.facing).value if with_normals and self.facing is not None else None)


def render_gbuffer(mesh: DeviceMesh, mvp: np.ndarray, width: int, height: int, stream=None) -> GBuffer:
    """Camera G-buffer (renderGBuffer, GBuffer.vert/.frag)."""
    dev = mesh.positions.device
    pos = torch.empty((3, height, width), dtype=torch.float32, device=dev)
    nrm = torch.empty((3, height, width), dtype=torch.float32, device=dev)
    fac = torch.empty((height, width), dtype=torch.float32, device=dev)
    nbytes = _lib.lib().gi_gbuffer_pass_workspace_size(mesh.ntris, width, height)
    ws = _WS.get(nbytes, dev)
    _lib.call("gi_gbuffer_pass", _dptr(mesh.positions), _dptr(mesh.normals), mesh.nverts, _dptr(mesh.indices),
              mesh.ntris, _f(mvp, 16), width, height, _dptr(pos), _dptr(nrm), _dptr(fac), _dptr(ws),
              ctypes.c_size_t(ws.numel()), _stream_handle(stream))
    return GBuffer(pos, nrm, fac)


def compute_hard_shadows(gbuf: GBuffer, depth: torch.Tensor, params: ShadowParams,
                         out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Hard-shadow screen pass (computeHardShadows); returns the R channel."""
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=gbuf.position.device)
    g = gbuf.to_c(with_normals=params.preEvaluation)
    p = params.to_c()
    ws = _WS.get(_lib.lib().gi_hard_shadow_workspace_size(gbuf.width, gbuf.height), gbuf.position.device)
    _lib.call("gi_hard_shadow_pass", ctypes.byref(g), _dptr(depth), ctypes.byref(p), _dptr(out), _dptr(ws),
              ctypes.c_size_t(ws.numel()), _stream_handle(stream))
    return out


@dataclass
class PCSSParams:
    """The ShadowParams fields PlausibleSoftShadow.frag reads with PCSS == 1
    (SoftShadowMapping/src/main.cpp:1604-1614 defaults).  `zNear` is the int
    uniform; the reference app never uploads it for this shader, so its value
    there is 0 (penumbra 0, every tested pixel lit); 1 gives the intended
    PCSS with MyGLGeometryViewer's near plane."""
    lightMVP: np.ndarray                 # unbiased P*V*M of the light centre
    shadowMapWidth: int
    shadowMapHeight: int
    blockerSearchSize: int = 7
    kernelSize: int = 15
    lightSourceRadius: int = 8           # LightSource::size / 2
    zNear: int = 1
    shadowIntensity: float = 0.25
    zbias: float = 0.0
    preEvaluation: bool = False
    MV: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32).reshape(-1))
    normalMatrix: np.ndarray = field(default_factory=lambda: np.eye(3, dtype=np.float32).reshape(-1))
    lightPosition: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> _lib.GiPcssParams:
        p = _lib.GiPcssParams()
        p.light_mvp = _f(biased(self.lightMVP, self.zbias), 16)
        p.shadow_map_width = int(self.shadowMapWidth)
        p.shadow_map_height = int(self.shadowMapHeight)
        p.blocker_search_size = int(self.blockerSearchSize)
        p.kernel_size = int(self.kernelSize)
        p.light_source_radius = int(self.lightSourceRadius)
        p.z_near = int(self.zNear)
        p.shadow_intensity = float(self.shadowIntensity)
        p.pre_evaluation = 1 if self.preEvaluation else 0
        p.mv = _f(self.MV, 16)
        p.normal_matrix = _f(self.normalMatrix, 9)
        p.light_position = _f(self.lightPosition, 3)
        return p


def render_soft_shadows(gbuf: GBuffer, depth: torch.Tensor, params: PCSSParams, out: torch.Tensor | None = None,
                        stream=None) -> torch.Tensor:
    """renderSoftShadows' screen pass with PCSS (PlausibleSoftShadow.frag):
    blocker search, penumbra width, PCF over one depth map of the light
    centre; returns the R channel."""
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=gbuf.position.device)
    g = gbuf.to_c(with_normals=params.preEvaluation)
    p = params.to_c()
    _lib.call("gi_pcss_pass", ctypes.byref(g), _dptr(depth), ctypes.byref(p), _dptr(out), _stream_handle(stream))
    return out


@dataclass
class RBSSMParams(PCSSParams):
    """The ShadowParams fields RBSSM.frag reads (SoftShadowMapping/src/main.cpp:
    1452-1454 selects it with kernelSize 7; maxSearch 16 at :1617, the
    depth threshold is the scene's `d` line).  As with PCSS, the app never
    uploads zNear, so its value there is 0 and every tested pixel is lit."""
    kernelSize: int = 7
    maxSearch: int = 16
    depthThreshold: float = 0.0

    def to_c(self) -> _lib.GiRbssmParams:
        p = _lib.GiRbssmParams()
        p.pcss = PCSSParams.to_c(self)
        p.max_search = int(self.maxSearch)
        p.depth_threshold = float(self.depthThreshold)
        return p


def render_rbssm(gbuf: GBuffer, depth: torch.Tensor, params: RBSSMParams, out: torch.Tensor | None = None,
                 stream=None) -> torch.Tensor:
    """renderSoftShadows' screen pass with RBSSM (RBSSM.frag): PCSS's blocker
    search and penumbra width, then a kernelSize^2 kernel whose taps on a
    shadow-map discontinuity are revectorized; returns the R channel."""
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=gbuf.position.device)
    g = gbuf.to_c(with_normals=params.preEvaluation)
    p = params.to_c()
    ws = _WS.get(_lib.lib().gi_rbssm_workspace_size(gbuf.width, gbuf.height), gbuf.position.device)
    _lib.call("gi_rbssm_pass", ctypes.byref(g), _dptr(depth), ctypes.byref(p), _dptr(out), _dptr(ws),
              ctypes.c_size_t(ws.numel()), _stream_handle(stream))
    return out


# G-buffer colour of a mesh without per-vertex colours: GBuffer.frag:27 reads
# gl_FrontLightModelProduct.sceneColor, the default material ambient (0.2) times
# the light-model ambient (0.1, MyGLGeometryViewer.cpp:69), alpha 1.
SCENE_COLOR = tuple(float(np.float32(0.2) * np.float32(0.1)) for _ in range(3)) + (1.0,)
CLEAR_COLOR = (0.63, 0.82, 0.96, 1.0)  # shadeScene / deferredShading, main.cpp:452


@dataclass
class PhongParams:
    """Uniforms of PhongShading.frag:1-8 as displaySceneFromGBuffer sets them."""
    MV: np.ndarray
    normalMatrix: np.ndarray
    lightPosition: tuple
    shadowIntensity: float = 0.25
    color: tuple = SCENE_COLOR
    clearColor: tuple = CLEAR_COLOR

    @classmethod
    def from_views(cls, cam: View, light_eye, shadowIntensity: float = 0.25, **kw) -> "PhongParams":
        MV, nm, lpos = phong_uniforms(cam, light_eye)
        return cls(MV, nm, lpos, shadowIntensity, **kw)

    def to_c(self) -> _lib.GiPhongParams:
        p = _lib.GiPhongParams()
        p.mv = _f(self.MV, 16)
        p.normal_matrix = _f(self.normalMatrix, 9)
        p.light_position = _f(self.lightPosition, 3)
        p.shadow_intensity = float(self.shadowIntensity)
        p.color = _f(self.color, 4)
        p.clear_color = _f(self.clearColor, 4)
        return p


def shade_scene(gbuf: GBuffer, shadow: torch.Tensor, params: PhongParams, color: torch.Tensor | None = None,
                out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Deferred Phong shading (shadeScene main.cpp:449-456 / deferredShading
    Soft main.cpp:1109-1117, PhongShading.frag): (H, W, 4) RGBA from the
    G-buffer and a hard or soft shadow buffer; `color` is an optional
    (H, W, 4) G-buffer colour plane."""
    if gbuf.facing is None:
        raise ValueError("shade_scene needs a G-buffer with its facing plane (render_gbuffer makes one)")
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width, 4), dtype=torch.float32, device=gbuf.position.device)
    g = gbuf.to_c(with_normals=True)
    p = params.to_c()
    _lib.call("gi_phong_pass", ctypes.byref(g), _dptr(shadow), _dptr(color) if color is not None else None,
              ctypes.byref(p), _dptr(out), _stream_handle(stream))
    return out


@dataclass
class SoftParams:
    shadowIntensity: float = 0.25
    pcf: bool = False
    kernelOrder: int = 3
    penumbra: float = 1.0
    inclusive: bool = True  # the reference loop (AccurateSoftShadow.frag:108-109)

    def to_c(self) -> _lib.GiSoftParams:
        return _lib.GiSoftParams(float(self.shadowIntensity), 1 if self.pcf else 0, int(self.kernelOrder),
                                 float(self.penumbra), 1 if self.inclusive else 0)


def soft_uniforms(light_mvps: np.ndarray, zbias: float = 0.0):
    """(common B*lightMVP, lightMVPTrans) of configureShadow
    (SoftShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:168-195): columns 0-2
    of the LAST sample's biased matrix, column 3 of every sample's."""
    m = np.ascontiguousarray(light_mvps, dtype=np.float32).reshape(-1, 16)
    b = np.stack([biased(x, zbias) for x in m])
    return b[-1].copy(), camera.soft_light_trans(b)


def render_monte_carlo(gbuf: GBuffer, maps: torch.Tensor, light_mvps: np.ndarray, params: SoftParams,
                       zbias: float = 0.0, out: torch.Tensor | None = None, trans_dev: torch.Tensor | None = None,
                       stream=None) -> torch.Tensor:
    """Area-light screen pass of renderMonteCarlo (AccurateSoftShadow.frag)."""
    S, smh, smw = (int(x) for x in maps.shape)
    common, trans = soft_uniforms(light_mvps, zbias)
    if trans.shape[0] != S:
        raise ValueError("one light matrix per depth map is required")
    if trans_dev is None:
        trans_dev = torch.from_numpy(trans).to(maps.device)
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=maps.device)
    g = gbuf.to_c(with_normals=False)
    p = params.to_c()
    _lib.call("gi_soft_shadow_pass", ctypes.byref(g), _dptr(maps), S, smw, smh, _f(common, 16),
              _dptr(trans_dev), ctypes.byref(p), _dptr(out), _stream_handle(stream))
    return out


def render_monte_carlo_chunk(gbuf: GBuffer, maps: torch.Tensor, common: np.ndarray, trans_dev: torch.Tensor,
                             params: SoftParams, acc: torch.Tensor, first: bool, total_samples: int,
                             out: torch.Tensor, stream=None) -> None:
    """One run of consecutive samples of a renderMonteCarlo frame
    (gi_soft_shadow_pass_chunk): `maps` (n, H, W) and `trans_dev` (n, 4) are
    the run's slices, `common` the frame's soft uniform (soft_uniforms), `acc`
    an (H, W) float32 running sum; total_samples > 0 on the last run."""
    n, smh, smw = (int(x) for x in maps.shape)
    if trans_dev.shape[0] != n:
        raise ValueError("one trans row per depth map of the run is required")
    g = gbuf.to_c(with_normals=False)
    p = params.to_c()
    _lib.call("gi_soft_shadow_pass_chunk", ctypes.byref(g), _dptr(maps), n, smw, smh, _f(common, 16),
              _dptr(trans_dev), ctypes.byref(p), _dptr(acc), 1 if first else 0, int(total_samples), _dptr(out),
              _stream_handle(stream))


def adaptive_trans(light_mvps: np.ndarray, levels, zbias: float = 0.0) -> tuple:
    """configureShadow for adaptive sampling (MyGLGeometryViewer.cpp:184-194):
    the soft uniforms with each sample's quad-tree level added to w as
    level * 10000 (FP32)."""
    common, trans = soft_uniforms(light_mvps, zbias)
    lv = np.asarray(levels, np.float32).reshape(-1)
    if lv.shape[0] != trans.shape[0]:
        raise ValueError("one level per light sample is required")
    trans = trans.copy()
    trans[:, 3] = trans[:, 3] + lv * np.float32(10000.0)
    return common, trans


def render_adaptive(gbuf: GBuffer, maps: torch.Tensor, common: np.ndarray, trans: np.ndarray, params: SoftParams,
                    out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Final pass of renderAdaptiveLightSourceSampling: AccurateSoftShadow.frag
    with adaptiveSampling == 1 (`common`, `trans` from adaptive_trans)."""
    S, smh, smw = (int(x) for x in maps.shape)
    trans = np.ascontiguousarray(trans, dtype=np.float32).reshape(-1, 4)
    if trans.shape[0] != S:
        raise ValueError("one trans row per depth map is required")
    trans_dev = torch.from_numpy(trans).to(maps.device)
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=maps.device)
    g = gbuf.to_c(with_normals=False)
    p = params.to_c()
    _lib.call("gi_adaptive_soft_shadow_pass", ctypes.byref(g), _dptr(maps), S, smw, smh, _f(common, 16),
              _dptr(trans_dev), ctypes.byref(p), _dptr(out), _stream_handle(stream))
    return out


def quadtree_node_test(gbuf: GBuffer, maps: torch.Tensor, map_index, biased_mvps: np.ndarray, div: int = 1,
                       image: torch.Tensor | None = None, stream=None) -> tuple:
    """One quad-tree node's test (QuadTree/Reprojection.frag +
    Evaluation.frag): (reprojection image, passing pixel count).  Reads the
    count back, as the reference reads its occlusion query."""
    S, smh, smw = (int(x) for x in maps.shape)
    if image is None:
        image = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=maps.device)
    count = torch.empty((1,), dtype=torch.int32, device=maps.device)
    idx = (ctypes.c_int * 4)(*[int(i) for i in map_index])
    g = gbuf.to_c(with_normals=False)
    _lib.call("gi_quadtree_node_test", ctypes.byref(g), _dptr(maps), S, smw, smh, idx, _f(biased_mvps, 64), int(div),
              _dptr(image), _dptr(count), _stream_handle(stream))
    return image, int(count.item())


# ---------------------------------------------------------------------------
# Revectorization-based adaptive sampling (ShadowParams.
# revectorizationBasedAdaptiveSampling, SoftShadowMapping/src/main.cpp:
# 440-571, 813-892)
# ---------------------------------------------------------------------------
def rev_level_div(level: int, rpcf: bool) -> int:
    """Evaluation grid divisor of a node test (main.cpp:535-549)."""
    if level <= 1:
        return 1
    if level <= 3:
        return 3 if rpcf else 2
    return 6 if rpcf else 4


def rev_discontinuity_layers(gbuf: GBuffer, layers: int = _lib.GI_REV_DISC_LAYERS) -> torch.Tensor:
    """The discontinuity array: `layers` x H x W bytes, zero-filled (the
    reference's layers start undefined and keep earlier frames' contents)."""
    return torch.zeros((layers, gbuf.height, gbuf.width), dtype=torch.uint8, device=gbuf.position.device)


def rev_discontinuity(gbuf: GBuffer, maps: torch.Tensor, map_index: int, biased_mvp: np.ndarray, MV: np.ndarray,
                      normalMatrix: np.ndarray, lightPosition, shadowIntensity: float, disc: torch.Tensor,
                      stream=None) -> None:
    """Discontinuity.frag for corner map `map_index` into layer map_index of
    `disc` (nothing is written past the array's last layer)."""
    S, smh, smw = (int(x) for x in maps.shape)
    g = gbuf.to_c(with_normals=True)
    _lib.call("gi_rev_discontinuity_pass", ctypes.byref(g), _dptr(maps), S, smw, smh, int(map_index),
              _f(biased_mvp, 16), _f(MV, 16), _f(normalMatrix, 9), _f(lightPosition, 3), float(shadowIntensity),
              _dptr(disc), int(disc.shape[0]), _stream_handle(stream))


def rev_quadtree_node_test(gbuf: GBuffer, maps: torch.Tensor, map_index, biased_mvps: np.ndarray,
                           disc: torch.Tensor, level: int, visibility: torch.Tensor, div: int = 1,
                           image: torch.Tensor | None = None, stream=None) -> tuple:
    """RevectorizationBasedReprojection.frag + Evaluation.frag: (image, pass
    count); `visibility` (H, W) float32 is updated in place."""
    S, smh, smw = (int(x) for x in maps.shape)
    if image is None:
        image = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=maps.device)
    count = torch.empty((1,), dtype=torch.int32, device=maps.device)
    idx = (ctypes.c_int * 4)(*[int(i) for i in map_index])
    g = gbuf.to_c(with_normals=False)
    _lib.call("gi_rev_quadtree_node_test", ctypes.byref(g), _dptr(maps), S, smw, smh, idx, _f(biased_mvps, 64),
              _dptr(disc), int(disc.shape[0]), int(level), int(div), _dptr(image), _dptr(visibility), _dptr(count),
              _stream_handle(stream))
    return image, int(count.item())


@dataclass
class RevSoftParams:
    """Uniforms of RevectorizationBasedAccurateSoftShadow.frag."""
    MV: np.ndarray
    normalMatrix: np.ndarray
    lightPosition: tuple
    shadowIntensity: float = 0.25
    RPCF: bool = False
    maxSearch: int = 16
    depthThreshold: float = 0.0

    def to_c(self) -> _lib.GiRevSoftParams:
        p = _lib.GiRevSoftParams()
        p.shadow_intensity = float(self.shadowIntensity)
        p.rpcf = 1 if self.RPCF else 0
        p.max_search = int(self.maxSearch)
        p.depth_threshold = float(self.depthThreshold)
        p.mv = _f(self.MV, 16)
        p.normal_matrix = _f(self.normalMatrix, 9)
        p.light_position = _f(self.lightPosition, 3)
        return p


def render_rev_adaptive(gbuf: GBuffer, maps: torch.Tensor, common: np.ndarray, trans: np.ndarray,
                        disc: torch.Tensor, visibility: torch.Tensor, params: RevSoftParams,
                        out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """Final pass of the revectorization-based adaptive frame (`common`,
    `trans` from adaptive_trans)."""
    S, smh, smw = (int(x) for x in maps.shape)
    trans = np.ascontiguousarray(trans, dtype=np.float32).reshape(-1, 4)
    if trans.shape[0] != S:
        raise ValueError("one trans row per depth map is required")
    trans_dev = torch.from_numpy(trans).to(maps.device)
    if out is None:
        out = torch.empty((gbuf.height, gbuf.width), dtype=torch.float32, device=maps.device)
    g = gbuf.to_c(with_normals=True)
    p = params.to_c()
    ws = _WS.get(_lib.lib().gi_rev_soft_workspace_size(gbuf.width, gbuf.height), maps.device)
    _lib.call("gi_rev_adaptive_soft_shadow_pass", ctypes.byref(g), _dptr(maps), S, smw, smh, _f(common, 16),
              _dptr(trans_dev), _dptr(disc), int(disc.shape[0]), _dptr(visibility), ctypes.byref(p), _dptr(out),
              _dptr(ws), ctypes.c_size_t(ws.numel()), _stream_handle(stream))
    return out
