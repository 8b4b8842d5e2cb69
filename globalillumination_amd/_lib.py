"""ctypes binding of libgi.so (the C ABI declared in include/gi.h).

torch is imported first so that the HIP runtime torch ships is the one libgi.so
binds to: the kernels then run on torch's streams and in torch's memory.  There
is no CPU fallback: if the library is missing or fails to load, every pass
raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch  # noqa: F401  (loads the HIP runtime before libgi.so)

LIB_PATH = Path(__file__).resolve().parent / "libgi.so"

GI_OK = 0
GI_EINVAL = -1
GI_ENONFINITE = -2
GI_EDEVICE = -3
GI_ENOSPACE = -4
GI_ENOTSUP = -5
GI_EIO = -6

GI_NAIVE, GI_PCF, GI_RBSM, GI_RPCF, GI_RBSM_FILTERED, GI_RBSM_CONSERVATIVE, GI_RPCF_CONSERVATIVE = range(7)
GI_MAX_SAMPLES = 4096
GI_MAX_EDGE = 16384
GI_REV_DISC_LAYERS = 50  # the reference's discontinuity array depth (main.cpp:1635)

F16 = ctypes.c_float * 16
F9 = ctypes.c_float * 9
F3 = ctypes.c_float * 3


class GiRasterOpts(ctypes.Structure):
    _fields_ = [("polygon_offset_factor", ctypes.c_float),
                ("polygon_offset_units", ctypes.c_float),
                ("raster_begin_event", ctypes.c_void_p),
                ("raster_end_event", ctypes.c_void_p)]


class GiGBuffer(ctypes.Structure):
    _fields_ = [("position_dev", ctypes.c_void_p),
                ("normal_dev", ctypes.c_void_p),
                ("width", ctypes.c_int),
                ("height", ctypes.c_int),
                ("facing_dev", ctypes.c_void_p)]


class GiSceneInfo(ctypes.Structure):
    _fields_ = [("nverts", ctypes.c_int), ("ntris", ctypes.c_int), ("has_colors", ctypes.c_int),
                ("ntextures", ctypes.c_int), ("camera_eye", ctypes.c_float * 3), ("camera_at", ctypes.c_float * 3),
                ("light_eye", ctypes.c_float * 3), ("light_at", ctypes.c_float * 3),
                ("depth_threshold", ctypes.c_float), ("hsm_alpha", ctypes.c_float), ("hsm_beta", ctypes.c_float)]


class GiPcssParams(ctypes.Structure):
    _fields_ = [("light_mvp", ctypes.c_float * 16),
                ("shadow_map_width", ctypes.c_int),
                ("shadow_map_height", ctypes.c_int),
                ("blocker_search_size", ctypes.c_int),
                ("kernel_size", ctypes.c_int),
                ("light_source_radius", ctypes.c_int),
                ("z_near", ctypes.c_int),
                ("shadow_intensity", ctypes.c_float),
                ("pre_evaluation", ctypes.c_int),
                ("mv", ctypes.c_float * 16),
                ("normal_matrix", ctypes.c_float * 9),
                ("light_position", ctypes.c_float * 3)]


class GiRbssmParams(ctypes.Structure):
    _fields_ = [("pcss", GiPcssParams),
                ("max_search", ctypes.c_int),
                ("depth_threshold", ctypes.c_float)]


class GiRevSoftParams(ctypes.Structure):
    _fields_ = [("shadow_intensity", ctypes.c_float),
                ("rpcf", ctypes.c_int),
                ("max_search", ctypes.c_int),
                ("depth_threshold", ctypes.c_float),
                ("mv", ctypes.c_float * 16),
                ("normal_matrix", ctypes.c_float * 9),
                ("light_position", ctypes.c_float * 3)]


class GiPhongParams(ctypes.Structure):
    _fields_ = [("mv", ctypes.c_float * 16),
                ("normal_matrix", ctypes.c_float * 9),
                ("light_position", ctypes.c_float * 3),
                ("shadow_intensity", ctypes.c_float),
                ("color", ctypes.c_float * 4),
                ("clear_color", ctypes.c_float * 4)]


class GiShadowParams(ctypes.Structure):
    _fields_ = [("light_mvp", F16),
                ("shadow_map_width", ctypes.c_int),
                ("shadow_map_height", ctypes.c_int),
                ("technique", ctypes.c_int),
                ("kernel_order", ctypes.c_int),
                ("penumbra_size", ctypes.c_int),
                ("max_search", ctypes.c_int),
                ("depth_threshold", ctypes.c_float),
                ("shadow_intensity", ctypes.c_float),
                ("pre_evaluation", ctypes.c_int),
                ("mv", F16),
                ("normal_matrix", F9),
                ("light_position", F3)]


class GiSoftParams(ctypes.Structure):
    _fields_ = [("shadow_intensity", ctypes.c_float),
                ("pcf", ctypes.c_int),
                ("kernel_order", ctypes.c_int),
                ("penumbra", ctypes.c_float),
                ("inclusive", ctypes.c_int)]


class GiDiffParams(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double),
                ("bias", ctypes.c_double),
                ("shadow_intensity", ctypes.c_double)]


class GiBatchInstance(ctypes.Structure):
    _fields_ = [("positions_dev", ctypes.c_void_p),
                ("nverts", ctypes.c_int),
                ("indices_dev", ctypes.c_void_p),
                ("ntris", ctypes.c_int),
                ("light_mvps_dev", ctypes.c_void_p),
                ("nsamples", ctypes.c_int),
                ("sm_width", ctypes.c_int),
                ("sm_height", ctypes.c_int),
                ("polygon_offset_factor", ctypes.c_float),
                ("polygon_offset_units", ctypes.c_float),
                ("maps_dev", ctypes.c_void_p),
                ("gbuffer", GiGBuffer),
                ("light_mvp", F16),
                ("trans_dev", ctypes.c_void_p),
                ("soft", GiSoftParams),
                ("out_dev", ctypes.c_void_p)]


class GiBatchTuning(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int),
                ("raster_groups", ctypes.c_int),
                ("screen_shape", ctypes.c_int),
                ("screen_groups", ctypes.c_int)]


class GiError(RuntimeError):
    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed: {msg} ({status})")
        self.status = status


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "gi_api_version": (ctypes.c_int, []),
    "gi_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "gi_perspective": (ctypes.c_int, [ctypes.c_float] * 4 + [ctypes.c_void_p]),
    "gi_ortho": (ctypes.c_int, [ctypes.c_float] * 6 + [ctypes.c_void_p]),
    "gi_lookat": (ctypes.c_int, [ctypes.c_void_p] * 4),
    "gi_mat4_mul": (ctypes.c_int, [ctypes.c_void_p] * 3),
    "gi_bias_matrix": (ctypes.c_int, [ctypes.c_float, ctypes.c_void_p]),
    "gi_uniform_sampling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p]),
    "gi_soft_light_trans": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "gi_depth_pass_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 4),
    "gi_depth_pass": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(GiRasterOpts), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_size_t, ctypes.c_void_p]),
    "gi_depth_pass_noalloc_max_bytes": (ctypes.c_size_t, []),
    "gi_depth_pass_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(GiRasterOpts), ctypes.c_void_p, ctypes.c_void_p]),
    "gi_gbuffer_pass_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 3),
    "gi_gbuffer_pass": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_void_p]),
    "gi_hard_shadow_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 2),
    "gi_hard_shadow_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p,
                                           ctypes.POINTER(GiShadowParams), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_size_t, ctypes.c_void_p]),
    "gi_depth_pass_bin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(GiRasterOpts), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_size_t, ctypes.c_void_p]),
    "gi_depth_pass_raster": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(GiRasterOpts), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "gi_soft_shadow_pass_chunk": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.POINTER(GiSoftParams), ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gi_soft_shadow_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(GiSoftParams), ctypes.c_void_p, ctypes.c_void_p]),
    "gi_scene_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "gi_obj_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "gi_scene_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GiSceneInfo)]),
    "gi_scene_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "gi_scene_texture_path": (ctypes.c_char_p, [ctypes.c_void_p, ctypes.c_int]),
    "gi_scene_free": (None, [ctypes.c_void_p]),
    "gi_quadtree_node_test": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float),
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "gi_adaptive_soft_shadow_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.POINTER(GiSoftParams), ctypes.c_void_p, ctypes.c_void_p]),
    "gi_pcss_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.POINTER(GiPcssParams),
                                    ctypes.c_void_p, ctypes.c_void_p]),
    "gi_rbssm_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 2),
    "gi_rbssm_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.POINTER(GiRbssmParams),
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "gi_rev_discontinuity_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "gi_rev_quadtree_node_test": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]),
    "gi_rev_adaptive_soft_shadow_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_int,
                                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                        ctypes.c_void_p, ctypes.POINTER(GiRevSoftParams),
                                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                        ctypes.c_void_p]),
    "gi_rev_soft_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 2),
    "gi_phong_pass": (ctypes.c_int, [ctypes.POINTER(GiGBuffer), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(GiPhongParams), ctypes.c_void_p, ctypes.c_void_p]),
    "gi_diff_workspace_size": (ctypes.c_size_t, [ctypes.c_int] * 6),
    "gi_diff_soft_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(GiGBuffer), ctypes.POINTER(GiDiffParams),
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "gi_diff_soft_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(GiGBuffer), ctypes.POINTER(GiDiffParams),
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "gi_batch_workspace_size": (ctypes.c_size_t, [ctypes.POINTER(GiBatchInstance), ctypes.c_int]),
    "gi_batch_create": (ctypes.c_int, [ctypes.POINTER(GiBatchInstance), ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "gi_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GiBatchTuning)]),
    "gi_batch_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(GiBatchTuning)]),
    "gi_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gi_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.POINTER(GiBatchTuning), ctypes.POINTER(ctypes.c_float)]),
    "gi_batch_graph_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "gi_batch_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gi_batch_graph_destroy": (None, [ctypes.c_void_p]),
    "gi_batch_destroy": (None, [ctypes.c_void_p]),
}

EXPORTED = tuple(_SIGS)


def lib() -> ctypes.CDLL:
    """Loads libgi.so once; raises when it is absent (no fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(fn_name: str, status: int) -> None:
    if status != GI_OK:
        msg = lib().gi_status_string(status).decode()
        raise GiError(fn_name, status, msg)


def call(fn_name: str, *args) -> None:
    check(fn_name, getattr(lib(), fn_name)(*args))
