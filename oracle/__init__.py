"""ctypes binding of the CPU oracle (oracle/libgi_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP path.  The product
package (globalillumination_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libgi_oracle.so"

NAIVE, PCF, RBSM, RPCF, RBSM_FILTERED, RBSM_CONSERVATIVE, RPCF_CONSERVATIVE = range(7)
TECHNIQUES = {"naive": NAIVE, "pcf": PCF, "rbsm": RBSM, "rpcf": RPCF, "rbsm_filtered": RBSM_FILTERED,
              "rbsm_conservative": RBSM_CONSERVATIVE, "rpcf_conservative": RPCF_CONSERVATIVE}

F16 = ctypes.c_float * 16


class OrParams(ctypes.Structure):
    _fields_ = [("light_mvp", F16), ("sm_width", ctypes.c_int), ("sm_height", ctypes.c_int),
                ("technique", ctypes.c_int), ("kernel_order", ctypes.c_int), ("penumbra_size", ctypes.c_int),
                ("max_search", ctypes.c_int), ("depth_threshold", ctypes.c_float),
                ("shadow_intensity", ctypes.c_float), ("pre_evaluation", ctypes.c_int), ("mv", F16),
                ("normal_matrix", ctypes.c_float * 9), ("light_position", ctypes.c_float * 3)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        sig = {
            "or_perspective": [f, f, f, f, vp], "or_ortho": [f] * 6 + [vp], "or_lookat": [vp] * 4,
            "or_mat4_mul": [vp] * 3, "or_bias_matrix": [f, vp], "or_uniform_sampling": [vp, i, i, i, vp],
            "or_fetch": [vp, i, i, f, f],
            "or_hard_shadow": [vp, vp, i, i, vp, ctypes.POINTER(OrParams), vp],
            "or_soft_shadow": [vp, i, i, vp, i, i, i, vp, vp, f, i, i, f, i, vp],
            "or_depth_raster": [vp, i, vp, i, vp, i, i, f, f, vp],
            "or_depth_raster_brute": [vp, i, vp, i, vp, i, i, f, f, vp],
            "or_visibility_raster": [vp, i, vp, i, vp, i, i, vp],
            "or_gbuffer": [vp, vp, i, vp, i, vp, i, i, vp, vp, vp],
            "or_phong": [vp, vp, vp, i, i, vp, vp, vp, vp],
            "or_pcss": [vp, vp, i, i, vp, vp, i, i, i, i, vp],
            "or_rbssm": [vp, vp, i, i, vp, vp, i, i, i, i, vp],
            "or_rbssm_tap": [vp, vp, f, f, f, vp],
            "or_rbsm_tap": [vp, vp, i, f, f, f, vp],
            "or_soft_shadow_adaptive": [vp, i, i, vp, i, i, i, vp, vp, ctypes.c_float, i, i, ctypes.c_float, i, vp],
            "or_qt_node_test": [vp, i, i, vp, i, i, vp, vp, i, vp, vp],
            "or_rev_discontinuity": [vp, vp, i, i, vp, vp, vp],
            "or_rev_node_test": [vp, i, i, vp, i, i, vp, vp, vp, i, i, i, vp, vp, vp],
            "or_rev_soft": [vp, vp, i, i, vp, i, i, i, vp, vp, vp, i, vp, vp, i, vp],
            "or_diff_depth": [vp, i, vp, i, vp, i, i, vp, vp],
            "or_diff_soft": [vp, i, i, vp, i, i, i, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp],
            "or_diff_adjoint": [vp, i, vp, i, vp, vp, vp, i, i, i, vp, i, i, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, vp, vp, vp, vp],
            "or_set_threads": [i], "or_get_threads": [],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_float if name == "or_fetch" else (None if name in (
                "or_perspective", "or_ortho", "or_lookat", "or_mat4_mul", "or_bias_matrix",
                "or_uniform_sampling", "or_set_threads") else ctypes.c_int)
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _c32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def threads() -> int:
    return int(lib().or_get_threads())


def perspective(fovy, aspect, zn, zf):
    o = np.zeros(16, np.float32)
    lib().or_perspective(fovy, aspect, zn, zf, _p(o))
    return o


def ortho(l, r, b, t, zn, zf):
    o = np.zeros(16, np.float32)
    lib().or_ortho(l, r, b, t, zn, zf, _p(o))
    return o


def lookat(eye, at, up):
    e, a, u = _c32(eye), _c32(at), _c32(up)
    o = np.zeros(16, np.float32)
    lib().or_lookat(_p(e), _p(a), _p(u), _p(o))
    return o


def mat4_mul(a, b):
    a, b = _c32(a), _c32(b)
    o = np.zeros(16, np.float32)
    lib().or_mat4_mul(_p(a), _p(b), _p(o))
    return o


def bias_matrix(zbias=0.0):
    o = np.zeros(16, np.float32)
    lib().or_bias_matrix(zbias, _p(o))
    return o


def uniform_sampling(base, size, n, index):
    b = _c32(base)
    o = np.zeros(3, np.float32)
    lib().or_uniform_sampling(_p(b), size, n, index, _p(o))
    return o


def fetch(depth: np.ndarray, s: float, t: float) -> float:
    d = _c32(depth)
    h, w = d.shape
    return float(lib().or_fetch(_p(d), w, h, s, t))


def make_params(light_mvp_biased, smw, smh, technique="naive", kernel_order=3, penumbra_size=1,
                max_search=16, depth_threshold=2.5e-6, shadow_intensity=0.25, pre_evaluation=False,
                mv=None, normal_matrix=None, light_position=(0, 0, 0)) -> OrParams:
    p = OrParams()
    p.light_mvp = F16(*_c32(light_mvp_biased).reshape(-1).tolist())
    p.sm_width, p.sm_height = int(smw), int(smh)
    p.technique = TECHNIQUES[technique] if isinstance(technique, str) else int(technique)
    p.kernel_order, p.penumbra_size, p.max_search = int(kernel_order), int(penumbra_size), int(max_search)
    p.depth_threshold, p.shadow_intensity = float(depth_threshold), float(shadow_intensity)
    p.pre_evaluation = 1 if pre_evaluation else 0
    mv = np.eye(4, dtype=np.float32).reshape(-1) if mv is None else _c32(mv).reshape(-1)
    nm = np.eye(3, dtype=np.float32).reshape(-1) if normal_matrix is None else _c32(normal_matrix).reshape(-1)
    p.mv = F16(*mv.tolist())
    p.normal_matrix = (ctypes.c_float * 9)(*nm.tolist())
    p.light_position = (ctypes.c_float * 3)(*[float(x) for x in light_position])
    return p


def hard_shadow(pos: np.ndarray, nrm: np.ndarray | None, depth: np.ndarray, params: OrParams) -> np.ndarray:
    pos = _c32(pos)
    _, h, w = pos.shape
    nrm_c = _c32(nrm) if nrm is not None else None
    d = _c32(depth)
    out = np.empty((h, w), np.float32)
    rc = lib().or_hard_shadow(_p(pos), _p(nrm_c) if nrm_c is not None else None, w, h, _p(d),
                              ctypes.byref(params), _p(out))
    if rc:
        raise RuntimeError(f"or_hard_shadow failed ({rc})")
    return out


def rbssm(pos, nrm, depth, params, blocker_search_size=7, kernel_size=7, light_source_radius=8, z_near=1):
    """RBSSM.frag over a G-buffer; `params` from make_params (light_mvp biased,
    map size, shadow intensity, pre-evaluation, max_search, depth_threshold)."""
    pos = _c32(pos)
    h, w = pos.shape[1], pos.shape[2]
    d = _c32(depth)
    assert d.shape == (params.sm_height, params.sm_width)
    n = None if nrm is None else _c32(nrm)
    out = np.empty((h, w), np.float32)
    rc = lib().or_rbssm(_p(pos), None if n is None else _p(n), w, h, _p(d), ctypes.byref(params),
                        int(blocker_search_size), int(kernel_size), int(light_source_radius), int(z_near), _p(out))
    if rc:
        raise RuntimeError(f"or_rbssm failed ({rc})")
    return out


def rbssm_tap(depth, params, x, y, z):
    """One RBSSM tap: dict of the discontinuity (r, g, b), the four edge-walk
    lengths, the normalized discontinuity and smoothONDS's fill."""
    d = _c32(depth)
    assert d.shape == (params.sm_height, params.sm_width)
    out = np.empty(10, np.float32)
    rc = lib().or_rbssm_tap(_p(d), ctypes.byref(params), float(x), float(y), float(z), _p(out))
    if rc:
        raise RuntimeError(f"or_rbssm_tap failed ({rc})")
    keys = ("r", "g", "b", "left", "right", "down", "up", "nx", "ny", "fill")
    return {k: float(v) for k, v in zip(keys, out)}


def rbsm_tap(depth, params, x, y, z, filtered=False):
    """One RBSM tap: dict of the discontinuity (r, g, b), the four edge-walk
    lengths, the two relative positions and revectorizeShadow's value (before
    the mix with the shadow intensity; the plain test without a
    discontinuity)."""
    d = _c32(depth)
    assert d.shape == (params.sm_height, params.sm_width)
    out = np.empty(10, np.float32)
    rc = lib().or_rbsm_tap(_p(d), ctypes.byref(params), int(bool(filtered)), float(x), float(y), float(z), _p(out))
    if rc:
        raise RuntimeError(f"or_rbsm_tap failed ({rc})")
    keys = ("r", "g", "b", "left", "right", "down", "up", "rx", "ry", "value")
    return {k: float(v) for k, v in zip(keys, out)}


def pcss(pos, nrm, depth, params, blocker_search_size=7, kernel_size=15, light_source_radius=8, z_near=1):
    """PlausibleSoftShadow.frag with PCSS == 1 over a G-buffer; `params` from
    make_params (light_mvp biased, map size, shadow intensity, pre-evaluation)."""
    pos = _c32(pos)
    h, w = pos.shape[1], pos.shape[2]
    d = _c32(depth)
    assert d.shape == (params.sm_height, params.sm_width)
    n = None if nrm is None else _c32(nrm)
    out = np.empty((h, w), np.float32)
    rc = lib().or_pcss(_p(pos), None if n is None else _p(n), w, h, _p(d), ctypes.byref(params),
                       int(blocker_search_size), int(kernel_size), int(light_source_radius), int(z_near), _p(out))
    if rc:
        raise RuntimeError(f"or_pcss failed ({rc})")
    return out


def soft_shadow(pos: np.ndarray, maps: np.ndarray, light_mvp_biased, trans, si=0.25, pcf=False,
                kernel_order=3, penumbra=1.0, inclusive=True) -> np.ndarray:
    pos, maps = _c32(pos), _c32(maps)
    _, h, w = pos.shape
    S, smh, smw = maps.shape
    m, t = _c32(light_mvp_biased).reshape(-1), _c32(trans).reshape(-1)
    out = np.empty((h, w), np.float32)
    rc = lib().or_soft_shadow(_p(pos), w, h, _p(maps), S, smw, smh, _p(m), _p(t), si, 1 if pcf else 0,
                              kernel_order, penumbra, 1 if inclusive else 0, _p(out))
    if rc:
        raise RuntimeError(f"or_soft_shadow failed ({rc})")
    return out


def soft_shadow_adaptive(pos: np.ndarray, maps: np.ndarray, light_mvp_biased, trans, si=0.25, pcf=False,
                         kernel_order=3, penumbra=1.0, inclusive=True) -> np.ndarray:
    """AccurateSoftShadow.frag with adaptiveSampling == 1; trans[:, 3] holds
    w + level * 10000 (float32)."""
    pos, maps = _c32(pos), _c32(maps)
    _, h, w = pos.shape
    S, smh, smw = maps.shape
    m, t = _c32(light_mvp_biased).reshape(-1), _c32(trans).reshape(-1)
    out = np.empty((h, w), np.float32)
    rc = lib().or_soft_shadow_adaptive(_p(pos), w, h, _p(maps), S, smw, smh, _p(m), _p(t), si, 1 if pcf else 0,
                                       kernel_order, penumbra, 1 if inclusive else 0, _p(out))
    if rc:
        raise RuntimeError(f"or_soft_shadow_adaptive failed ({rc})")
    return out


def qt_node_test(pos: np.ndarray, maps: np.ndarray, idx, mvps_biased, div=1):
    """One quad-tree node's reprojection image and pass count."""
    pos, maps = _c32(pos), _c32(maps)
    _, h, w = pos.shape
    S, smh, smw = maps.shape
    ix = np.ascontiguousarray(idx, dtype=np.int32).reshape(4)
    assert (ix >= 0).all() and (ix < S).all()
    m = _c32(mvps_biased).reshape(64)
    img = np.empty((h, w), np.float32)
    cnt = ctypes.c_uint(0)
    rc = lib().or_qt_node_test(_p(pos), w, h, _p(maps), smw, smh, _p(ix), _p(m), int(div), _p(img), ctypes.byref(cnt))
    if rc:
        raise RuntimeError(f"or_qt_node_test failed ({rc})")
    return img, int(cnt.value)


REV_DISC_LAYERS = 50  # the reference's discontinuity array depth


def rev_discontinuity(pos, nrm, depth, params) -> np.ndarray:
    """Discontinuity.frag for one map: (h, w) uint8 codes (bits 0-1 r*4,
    2-3 g*4, 4 discType, 5 alpha; 0xFF background).  `params` from
    make_params: light_mvp = B * the map's lightMVP, map size, max_search,
    depth_threshold, the pre-evaluation's mv / normal_matrix /
    light_position and shadow_intensity."""
    pos, nrm, d = _c32(pos), _c32(nrm), _c32(depth)
    _, h, w = pos.shape
    assert nrm.shape == pos.shape and d.shape == (params.sm_height, params.sm_width)
    out = np.empty((h, w), np.uint8)
    rc = lib().or_rev_discontinuity(_p(pos), _p(nrm), w, h, _p(d), ctypes.byref(params),
                                    out.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise RuntimeError(f"or_rev_discontinuity failed ({rc})")
    return out


def rev_node_test(pos, maps, idx, mvps_biased, disc, level, vis, div=1):
    """RevectorizationBasedReprojection.frag + Evaluation.frag.  `disc`:
    (L, h, w) uint8 layers; `vis` (h, w) float32 is updated in place.
    Returns (image, pass count)."""
    pos, maps = _c32(pos), _c32(maps)
    _, h, w = pos.shape
    S, smh, smw = maps.shape
    ix = np.ascontiguousarray(idx, dtype=np.int32).reshape(4)
    assert (ix >= 0).all() and (ix < S).all()
    disc = np.ascontiguousarray(disc, dtype=np.uint8)
    assert disc.ndim == 3 and disc.shape[1:] == (h, w)
    assert vis.dtype == np.float32 and vis.shape == (h, w) and vis.flags.c_contiguous
    m = _c32(mvps_biased).reshape(64)
    img = np.empty((h, w), np.float32)
    cnt = ctypes.c_uint(0)
    rc = lib().or_rev_node_test(_p(pos), w, h, _p(maps), smw, smh, _p(ix), _p(m),
                                disc.ctypes.data_as(ctypes.c_void_p), disc.shape[0], int(level), int(div), _p(img),
                                _p(vis), ctypes.byref(cnt))
    if rc:
        raise RuntimeError(f"or_rev_node_test failed ({rc})")
    return img, int(cnt.value)


def rev_soft(pos, nrm, maps, light_mvp_biased, trans, disc, vis, params, rpcf=False) -> np.ndarray:
    """RevectorizationBasedAccurateSoftShadow.frag: maps (N, smh, smw),
    trans (N, 4) with w + level * 10000, disc (L, h, w) uint8, vis (h, w);
    `params` supplies shadow_intensity, max_search, depth_threshold and the
    pre-evaluation fields."""
    pos, nrm, maps = _c32(pos), _c32(nrm), _c32(maps)
    _, h, w = pos.shape
    n, smh, smw = maps.shape
    tr = _c32(trans).reshape(n, 4)
    disc = np.ascontiguousarray(disc, dtype=np.uint8)
    assert disc.ndim == 3 and disc.shape[1:] == (h, w)
    vis = _c32(vis)
    m = _c32(light_mvp_biased).reshape(16)
    out = np.empty((h, w), np.float32)
    rc = lib().or_rev_soft(_p(pos), _p(nrm), w, h, _p(maps), n, smw, smh, _p(m), _p(tr),
                           disc.ctypes.data_as(ctypes.c_void_p), disc.shape[0], _p(vis), ctypes.byref(params),
                           1 if rpcf else 0, _p(out))
    if rc:
        raise RuntimeError(f"or_rev_soft failed ({rc})")
    return out


def _mesh_args(positions, indices):
    pos = _c32(positions).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    return pos, idx


def depth_raster(positions, indices, mvp, w, h, offset=(0.0, 0.0), brute=False) -> np.ndarray:
    pos, idx = _mesh_args(positions, indices)
    m = _c32(mvp).reshape(-1)
    out = np.empty((h, w), np.float32)
    fn = lib().or_depth_raster_brute if brute else lib().or_depth_raster
    rc = fn(_p(pos), pos.shape[0], _p(idx), idx.shape[0], _p(m), w, h, offset[0], offset[1], _p(out))
    if rc:
        raise RuntimeError(f"or_depth_raster failed ({rc})")
    return out


def visibility_raster(positions, indices, mvp, w, h) -> np.ndarray:
    pos, idx = _mesh_args(positions, indices)
    m = _c32(mvp).reshape(-1)
    out = np.empty((h, w), np.uint64)
    rc = lib().or_visibility_raster(_p(pos), pos.shape[0], _p(idx), idx.shape[0], _p(m), w, h, _p(out))
    if rc:
        raise RuntimeError(f"or_visibility_raster failed ({rc})")
    return out


def gbuffer(positions, normals, indices, mvp, w, h):
    pos, idx = _mesh_args(positions, indices)
    nrm = _c32(normals).reshape(-1, 3)
    m = _c32(mvp).reshape(-1)
    po = np.empty((3, h, w), np.float32)
    no = np.empty((3, h, w), np.float32)
    fo = np.empty((h, w), np.float32)
    rc = lib().or_gbuffer(_p(pos), _p(nrm), pos.shape[0], _p(idx), idx.shape[0], _p(m), w, h, _p(po), _p(no),
                          _p(fo))
    if rc:
        raise RuntimeError(f"or_gbuffer failed ({rc})")
    return po, no, fo


class _PhongParams(ctypes.Structure):
    _fields_ = [("mv", ctypes.c_float * 16), ("normal_matrix", ctypes.c_float * 9),
                ("light_position", ctypes.c_float * 3), ("shadow_intensity", ctypes.c_float),
                ("color", ctypes.c_float * 4), ("clear_color", ctypes.c_float * 4)]


# gl_FrontLightModelProduct.sceneColor for the default material: ambient 0.2
# times the light-model ambient 0.1 (ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:69)
_SCENE = float(np.float32(0.2) * np.float32(0.1))


def phong(position, normal, facing, shadow, mv, normal_matrix, light_position, shadow_intensity,
          color=(_SCENE, _SCENE, _SCENE, 1.0), clear_color=(0.63, 0.82, 0.96, 1.0), color_plane=None) -> np.ndarray:
    """PhongShading.frag over a G-buffer: (H, W, 4) RGBA.  `color` is the
    G-buffer colour of every pixel unless `color_plane` (H, W, 4) is given."""
    pos = _c32(position)
    h, w = pos.shape[1], pos.shape[2]
    nrm = _c32(normal)
    fac = _c32(facing)
    sh = _c32(shadow)
    assert pos.shape == (3, h, w) and nrm.shape == (3, h, w) and fac.shape == (h, w) and sh.shape == (h, w)
    cp = None if color_plane is None else _c32(color_plane)
    if cp is not None:
        assert cp.shape == (h, w, 4)
    p = _PhongParams()
    for name, val, n in (("mv", mv, 16), ("normal_matrix", normal_matrix, 9), ("light_position", light_position, 3),
                         ("color", color, 4), ("clear_color", clear_color, 4)):
        a = _c32(val).reshape(-1)
        assert a.size == n, name
        setattr(p, name, (ctypes.c_float * n)(*a.tolist()))
    p.shadow_intensity = float(np.float32(shadow_intensity))
    out = np.empty((h, w, 4), np.float32)
    rc = lib().or_phong(_p(pos), _p(nrm), _p(fac), w, h, _p(sh), None if cp is None else _p(cp), ctypes.byref(p),
                        _p(out))
    if rc:
        raise RuntimeError(f"or_phong failed ({rc})")
    return out


def _c64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def diff_depth(positions, indices, mvp64, vis: np.ndarray) -> np.ndarray:
    """Continuous-model texel depths (float64) of one map; vis from
    visibility_raster fixes the triangle of every texel."""
    pos = _c64(positions).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    m = _c64(mvp64).reshape(-1)
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    h, w = v.shape
    out = np.empty((h, w), np.float64)
    rc = lib().or_diff_depth(_p(pos), pos.shape[0], _p(idx), idx.shape[0], _p(m), w, h, _p(v), _p(out))
    if rc:
        raise RuntimeError(f"or_diff_depth failed ({rc})")
    return out


def diff_soft(gpos: np.ndarray, maps64: np.ndarray, biased_mvps64, tau, bias, si) -> np.ndarray:
    """Continuous-model soft visibility (float64) of every G-buffer pixel."""
    g = _c32(gpos)
    _, gh, gw = g.shape
    maps = _c64(maps64)
    S, h, w = maps.shape
    m = _c64(biased_mvps64).reshape(S, 16)
    out = np.empty((gh, gw), np.float64)
    rc = lib().or_diff_soft(_p(g), gw, gh, _p(maps), S, w, h, _p(m), float(tau), float(bias), float(si), _p(out))
    if rc:
        raise RuntimeError(f"or_diff_soft failed ({rc})")
    return out


def diff_adjoint(positions, indices, mvps64, biased_mvps64, vis: np.ndarray, gpos: np.ndarray, tau, bias, si,
                 grad_out) -> tuple:
    """Float64 reverse-mode derivative of diff_depth + diff_soft for
    L = sum(grad_out * S), the triangle of every texel held fixed by vis
    (S x h x w).  mvps64 / biased_mvps64: (S, 16) column-major.  Returns
    (dL/dpositions (nv, 3), dL/dM'_i (S, 16), dL/d(B*M'_i) (S, 16)), the
    matrices column-major."""
    pos = _c64(positions).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    v = np.ascontiguousarray(vis, dtype=np.uint64)
    S, h, w = v.shape
    m = _c64(mvps64).reshape(S, 16)
    mb = _c64(biased_mvps64).reshape(S, 16)
    g = _c32(gpos)
    _, gh, gw = g.shape
    go = _c64(grad_out).reshape(gh, gw)
    gp = np.empty_like(pos)
    gm = np.empty((S, 16), np.float64)
    gb = np.empty((S, 16), np.float64)
    rc = lib().or_diff_adjoint(_p(pos), pos.shape[0], _p(idx), idx.shape[0], _p(m), _p(mb), _p(v), S, w, h, _p(g),
                               gw, gh, float(tau), float(bias), float(si), _p(go), _p(gp), _p(gm), _p(gb))
    if rc:
        raise RuntimeError(f"or_diff_adjoint failed ({rc})")
    return gp, gm, gb
