"""The C ABI boundary (include/*.h): the library loads, exports every
declared entry point, and rejects bad arguments before touching the GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = "\n".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gi_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(gi):
    lib = gi.lib()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    from globalillumination_amd import _lib

    assert set(names) <= set(_lib.EXPORTED), set(names) - set(_lib.EXPORTED)


def test_status_strings(gi):
    lib = gi.lib()
    assert lib.gi_api_version() == 1
    for code, text in [(0, b"ok"), (-1, b"invalid argument"), (-2, b"non-finite parameter"),
                       (-4, b"workspace too small")]:
        assert lib.gi_status_string(code) == text


def _fake(n=1):
    # non-null pointers that are never dereferenced (validation fails first)
    return ctypes.c_void_p(0x1000 * n)


def _no_gpu():
    """The argument checks below pass fake device pointers; they run only where
    no GPU is visible, so a validation slip can never reach a kernel launch."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("fake-pointer validation checks run on CPU-only hosts")


def test_depth_pass_argument_checks(gi):
    lib = gi.lib()
    from globalillumination_amd import _lib

    _no_gpu()
    opts = _lib.GiRasterOpts(0.0, 0.0)
    ws = lib.gi_depth_pass_workspace_size(4, 2, 256, 256)
    assert ws > 0
    assert lib.gi_depth_pass_workspace_size(4, 4, 256, 256) > ws  # grows with the batch
    assert lib.gi_depth_pass_workspace_size(100, 2, 0, 256) == 0
    assert lib.gi_depth_pass_workspace_size(100, 2, 1 << 20, 256) == 0
    args = [_fake(1), 10, _fake(2), 4, _fake(3), 2, 256, 256, ctypes.byref(opts), _fake(4), _fake(5), ws, None]
    bad = list(args)
    bad[0] = None
    assert lib.gi_depth_pass(*bad) == _lib.GI_EINVAL
    bad = list(args)
    bad[5] = 0
    assert lib.gi_depth_pass(*bad) == _lib.GI_EINVAL
    bad = list(args)
    bad[6] = 20000
    assert lib.gi_depth_pass(*bad) == _lib.GI_ENOTSUP
    bad = list(args)
    bad[11] = ws - 1
    assert lib.gi_depth_pass(*bad) == _lib.GI_ENOSPACE
    bad = list(args)
    bad[9] = ctypes.c_void_p(0x1004)  # depth output must be 16-byte aligned
    assert lib.gi_depth_pass(*bad) == _lib.GI_EINVAL
    nan_opts = _lib.GiRasterOpts(float("nan"), 0.0)
    bad = list(args)
    bad[8] = ctypes.byref(nan_opts)
    assert lib.gi_depth_pass(*bad) == _lib.GI_ENONFINITE


def test_hard_pass_argument_checks(gi):
    lib = gi.lib()
    from globalillumination_amd import _lib

    _no_gpu()
    g = _lib.GiGBuffer(0x2000, None, 64, 32)
    p = _lib.GiShadowParams()
    p.light_mvp = (ctypes.c_float * 16)(*np.eye(4, dtype=np.float32).reshape(-1))
    p.shadow_map_width = p.shadow_map_height = 128
    p.technique = _lib.GI_PCF
    p.kernel_order, p.penumbra_size, p.max_search = 3, 0, 16
    p.shadow_intensity = 0.25
    # zero penumbra: the PCF loop would never advance
    assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), None, 0, None) == _lib.GI_EINVAL
    p.penumbra_size = 1
    p.technique = 9
    assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), None, 0, None) == _lib.GI_EINVAL
    p.technique = _lib.GI_NAIVE
    p.light_mvp[5] = float("inf")
    assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), None, 0, None) == _lib.GI_ENONFINITE
    p.light_mvp[5] = 1.0
    p.pre_evaluation = 1  # needs normals
    assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), None, 0, None) == _lib.GI_EINVAL
    assert lib.gi_hard_shadow_pass(None, _fake(), ctypes.byref(p), _fake(2), None, 0, None) == _lib.GI_EINVAL
    # RPCF, FilteredRBSM and conservative RPCF need the workspace: missing,
    # misaligned or short
    p.pre_evaluation = 0
    assert lib.gi_hard_shadow_workspace_size(64, 32) == 256 + 64 * 32 * 4
    assert lib.gi_hard_shadow_workspace_size(0, 32) == 0
    for tech in (_lib.GI_RPCF, _lib.GI_RBSM_FILTERED, _lib.GI_RPCF_CONSERVATIVE):
        p.technique = tech
        assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), None, 1 << 20,
                                       None) == _lib.GI_EINVAL
        assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), ctypes.c_void_p(0x3004),
                                       1 << 20, None) == _lib.GI_EINVAL
        assert lib.gi_hard_shadow_pass(ctypes.byref(g), _fake(), ctypes.byref(p), _fake(2), _fake(16),
                                       256 + 64 * 32 * 4 - 1, None) == _lib.GI_ENOSPACE


def test_soft_pass_argument_checks(gi):
    lib = gi.lib()
    from globalillumination_amd import _lib

    _no_gpu()
    g = _lib.GiGBuffer(0x2000, None, 64, 32)
    sp = _lib.GiSoftParams(0.25, 1, 3, 1.0, 1)
    m = (ctypes.c_float * 16)(*np.eye(4, dtype=np.float32).reshape(-1))
    ok_args = [ctypes.byref(g), _fake(), 4, 128, 128, m, ctypes.c_void_p(0x3000), ctypes.byref(sp), _fake(5), None]
    bad = list(ok_args)
    bad[2] = _lib.GI_MAX_SAMPLES + 1
    assert lib.gi_soft_shadow_pass(*bad) == _lib.GI_EINVAL
    bad = list(ok_args)
    bad[6] = ctypes.c_void_p(0x3004)  # trans must be float4-aligned
    assert lib.gi_soft_shadow_pass(*bad) == _lib.GI_EINVAL
    sp0 = _lib.GiSoftParams(0.25, 1, 3, 0.0, 1)
    bad = list(ok_args)
    bad[7] = ctypes.byref(sp0)
    assert lib.gi_soft_shadow_pass(*bad) == _lib.GI_EINVAL


def test_host_math_errors(gi):
    from globalillumination_amd import camera, _lib

    with pytest.raises(_lib.GiError):
        camera.perspective(0.0, 1.0, 1.0, 10.0)
    with pytest.raises(_lib.GiError):
        camera.lookat((0, 0, 0), (0, 0, 0), (0, 0, 1))  # eye == center
    with pytest.raises(_lib.GiError):
        camera.lookat((0, 0, 0), (0, 0, -1), (0, 0, 1))  # up parallel to the view axis
    with pytest.raises(_lib.GiError):
        camera.bias_matrix(float("nan"))


def test_scene_generators_are_deterministic(gi):
    from globalillumination_amd import scenes

    a = scenes.config3(nsamples=4, freq=12)
    b = scenes.config3(nsamples=4, freq=12)
    assert (a.mesh.positions == b.mesh.positions).all() and (a.mesh.indices == b.mesh.indices).all()
    assert [v.eye for v in a.lights] == [v.eye for v in b.lights]
    c1 = scenes.config1()
    assert c1.mesh.ntris == 64 * 30 * 2 + 2 * 64 + 2  # 3968 sphere triangles + the quad
    assert scenes.displaced_icosphere(71, seed=3).ntris == 20 * 71 * 71
    soup = scenes.triangle_soup(1000, seed=1)
    cen = soup.positions.reshape(-1, 3, 3).mean(axis=1)
    assert cen.min(axis=0)[1] >= 0.5 - 1e-3 and cen.max(axis=0)[1] <= 4.0 + 1e-3
    offs = scenes.stratified_jitter(64, seed=2)
    cells = np.floor((offs + 0.5) * 8).astype(int)
    assert len({tuple(c) for c in cells}) == 64  # one sample per stratum
