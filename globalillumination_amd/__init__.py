"""MI355X-native shadow pipeline (light-view depth, hard/PCF/RBSM shadows,
area-light soft shadows) behind the C ABI of include/gi.h.

The compute path is libgi.so (hand-written gfx950 HIP kernels); this package
is the host-side harness: deterministic scenes, camera math through the C
ABI, and pass functions that mirror the reference's per-technique passes.
"""
from . import _lib, camera, diff, scenes
from ._lib import GiError, lib
from .passes import (DeviceMesh, GBuffer, PCSSParams, PhongParams, RBSSMParams, RevSoftParams, ShadowParams,
                     SoftParams, biased,
                     compute_hard_shadows, phong_uniforms, render_gbuffer, render_monte_carlo, render_rbssm,
                     render_shadow_map, render_soft_shadows, shade_scene, soft_uniforms, view_matrices)

__all__ = ["_lib", "camera", "diff", "scenes", "GiError", "lib", "DeviceMesh", "GBuffer", "PCSSParams",
           "PhongParams", "RBSSMParams", "RevSoftParams", "ShadowParams", "SoftParams", "biased", "compute_hard_shadows", "phong_uniforms",
           "render_gbuffer", "render_monte_carlo", "render_rbssm", "render_shadow_map", "render_soft_shadows", "shade_scene",
           "soft_uniforms", "view_matrices"]
