"""Parity of the HIP passes (through the C ABI) against the CPU oracle.

Binary visibility, depth maps and G-buffers must be bit-identical; PCF, RBSM
and soft-shadow values must agree within 1e-5 absolute (they are bit-identical
as well, which is asserted first and reported on failure).
"""
import numpy as np
import pytest
import torch

from helpers import (bits_equal, camera_mvp, clip_heavy_mesh, clipped_count, light_mvps, mismatch_report, overflow_mesh,
                     shrink_soup, wide_overflow_mesh)

pytestmark = pytest.mark.gpu

TOL = 1e-5  # absolute tolerance for PCF / RBSM / soft-shadow values


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def _depth_both(gi, oracle_lib, mesh, mvps, w, h, offset=(0.0, 0.0)):
    dm = gi.DeviceMesh.from_mesh(mesh)
    got = _np(gi.render_shadow_map(dm, mvps, w, h, polygon_offset=offset))
    want = np.stack([oracle_lib.depth_raster(mesh.positions, mesh.indices, m, w, h, offset) for m in mvps])
    return got, want


def test_depth_pass_config1(gi, oracle_lib, cuda):
    wl = gi.scenes.config1()
    mvps = light_mvps(wl)
    got, want = _depth_both(gi, oracle_lib, wl.mesh, mvps, 1024, 1024)
    assert bits_equal(got, want), mismatch_report(got, want, "cfg1 depth")
    assert (want < 1.0).mean() > 0.5


def test_depth_pass_polygon_offset(gi, oracle_lib, cuda):
    wl = gi.scenes.config1()
    mvps = light_mvps(wl)
    got, want = _depth_both(gi, oracle_lib, wl.mesh, mvps, 1024, 1024, offset=(4.0, 20.0))
    assert bits_equal(got, want), mismatch_report(got, want, "cfg1 depth, offset (4,20)")


def test_depth_pass_soup_ortho(gi, oracle_lib, cuda):
    mesh = shrink_soup(20_000)
    wl = gi.scenes.config2()
    light = gi.scenes.ortho_fit((0.3, -1.0, 0.2), mesh, 512, 512)
    mvps = np.stack([gi.passes.view_matrices(light)[2]])
    got, want = _depth_both(gi, oracle_lib, mesh, mvps, 512, 512)
    assert bits_equal(got, want), mismatch_report(got, want, "soup depth")
    del wl


def test_depth_pass_batch_of_area_light_samples(gi, oracle_lib, cuda):
    wl = gi.scenes.config3(nsamples=4, sm=1024, freq=40)
    mvps = light_mvps(wl)
    got, want = _depth_both(gi, oracle_lib, wl.mesh, mvps, 1024, 1024)
    assert bits_equal(got, want), mismatch_report(got, want, "area-light depth batch")


def test_depth_pass_device_matrices(gi, cuda):
    """Matrices handed over as a device tensor give the same maps as host
    matrices; a wrongly shaped tensor is rejected."""
    wl = gi.scenes.config3(nsamples=4, view_w=64, view_h=64, sm=256, freq=20)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    mvps = light_mvps(wl)
    want = gi.render_shadow_map(dm, mvps, 256, 256)
    dev = torch.from_numpy(np.ascontiguousarray(mvps, dtype=np.float32)).cuda()
    got = gi.render_shadow_map(dm, dev, 256, 256)
    assert bits_equal(_np(got), _np(want))
    with pytest.raises(ValueError):
        gi.render_shadow_map(dm, dev[:, :15].contiguous(), 256, 256)


def test_depth_pass_split_matches_single_call(gi, cuda):
    """gi_depth_pass_bin, then gi_depth_pass_raster over uneven ranges of
    maps, gives gi_depth_pass's maps bit for bit; bad ranges are rejected."""
    wl = gi.scenes.config3(nsamples=9, view_w=64, view_h=64, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    mvps = light_mvps(wl)
    want = gi.render_shadow_map(dm, mvps, 1024, 1024, polygon_offset=(4.0, 20.0))
    sp = gi.passes.SplitDepthPass(dm, mvps, 1024, 1024, polygon_offset=(4.0, 20.0))
    sp.out.fill_(-1.0)
    sp.bin()
    for b, n in ((0, 2), (2, 5), (7, 1), (8, 1)):
        sp.raster(b, n)
    assert bits_equal(_np(sp.out), _np(want))
    for b, n in ((-1, 2), (8, 2), (0, 0)):
        with pytest.raises(gi.GiError):
            sp.raster(b, n)


@pytest.mark.parametrize("pcf", [False, True])
def test_soft_pass_runs_match_single_pass(gi, cuda, pcf):
    """gi_soft_shadow_pass_chunk over runs of 5, 7 and 4 samples carries the
    running sums and gives gi_soft_shadow_pass's output bit for bit."""
    wl = gi.scenes.config3(nsamples=16, view_w=640, view_h=360, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 640, 360)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 1024, 1024)
    params = gi.SoftParams(pcf=pcf, kernelOrder=3, penumbra=1.0, inclusive=False)
    want = _np(gi.render_monte_carlo(g, maps, mvps, params))
    common, trans = gi.soft_uniforms(mvps)
    trans_dev = torch.from_numpy(trans).cuda()
    acc = torch.full((360, 640), 123.0, device="cuda")
    out = torch.full((360, 640), -7.0, device="cuda")
    b = 0
    for n in (5, 7, 4):
        gi.passes.render_monte_carlo_chunk(g, maps[b:b + n], common, trans_dev[b:b + n], params, acc, b == 0,
                                           16 if b + n == 16 else 0, out)
        b += n
    assert bits_equal(_np(out), want)


def test_frame_on_two_streams_matches_sequential(gi, cuda):
    """The frame in runs of 4 maps: bin and rasters on one stream, each run's
    soft pass on a second stream after that run's raster; same bits as the
    single depth pass and soft pass."""
    wl = gi.scenes.config3(nsamples=16, view_w=640, view_h=360, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 640, 360)
    mvps = light_mvps(wl)
    params = gi.SoftParams(pcf=True, kernelOrder=3, penumbra=1.0, inclusive=False)
    want_maps = gi.render_shadow_map(dm, mvps, 1024, 1024)
    want = _np(gi.render_monte_carlo(g, want_maps, mvps, params))
    common, trans = gi.soft_uniforms(mvps)
    trans_dev = torch.from_numpy(trans).cuda()
    acc = torch.empty((360, 640), device="cuda")
    out = torch.empty((360, 640), device="cuda")
    sa, sb = torch.cuda.current_stream(), torch.cuda.Stream()
    sp = gi.passes.SplitDepthPass(dm, mvps, 1024, 1024)
    for rep in range(3):
        out.fill_(-1.0)
        sb.wait_stream(sa)
        sp.bin(stream=sa)
        for c in range(4):
            sp.raster(4 * c, 4, stream=sa)
            ev = torch.cuda.Event()
            ev.record(sa)
            sb.wait_event(ev)
            gi.passes.render_monte_carlo_chunk(g, sp.out[4 * c:4 * c + 4], common, trans_dev[4 * c:4 * c + 4], params,
                                               acc, c == 0, 16 if c == 3 else 0, out, stream=sb)
        sa.wait_stream(sb)
        torch.cuda.synchronize()
        assert bits_equal(_np(sp.out), _np(want_maps)), rep
        assert bits_equal(_np(out), want), rep


@pytest.mark.parametrize("w,h", [(1, 1), (33, 17), (100, 257)])
def test_depth_pass_ragged_sizes(gi, oracle_lib, cuda, w, h):
    wl = gi.scenes.config1()
    light = gi.scenes.View((4.0, 8.0, 4.0), (0.0, 0.0, 0.0), w, h)
    mvps = np.stack([gi.passes.view_matrices(light)[2]])
    got, want = _depth_both(gi, oracle_lib, wl.mesh, mvps, w, h)
    assert bits_equal(got, want), mismatch_report(got, want, f"{w}x{h} depth")


def test_depth_pass_empty_mesh(gi, cuda):
    mesh = gi.scenes.ground_quad(1.0, 1.0, 0.0)
    dm = gi.DeviceMesh.from_mesh(mesh)
    dm.indices = dm.indices[:0].contiguous()
    mvps = np.stack([gi.passes.view_matrices(gi.scenes.View((0, 5, 0.1), (0, 0, 0), 64, 64))[2]])
    got = _np(gi.render_shadow_map(dm, mvps, 64, 64))
    assert (got == 1.0).all()


def test_depth_pass_overflow_fallback(gi, oracle_lib, cuda):
    # 64 large overlapping triangles: every one touches every tile, so the tile
    # lists exceed the workspace bound and the raster scans the whole mesh.
    mesh, light = overflow_mesh()
    mvps = np.stack([gi.passes.view_matrices(light)[2]])
    got, want = _depth_both(gi, oracle_lib, mesh, mvps, 1024, 1024)
    assert bits_equal(got, want), mismatch_report(got, want, "overflow fallback")


def test_depth_pass_wide_record_overflow(gi, oracle_lib, cuda):
    # 3000 mid-sized triangles, each wider than the binning's per-thread span:
    # more setup records than the workspace holds, so the map falls back.
    mesh, light = wide_overflow_mesh()
    mvps = np.stack([gi.passes.view_matrices(light)[2]])
    got, want = _depth_both(gi, oracle_lib, mesh, mvps, 512, 512)
    assert bits_equal(got, want), mismatch_report(got, want, "wide record overflow")


def test_depth_pass_clip_heavy(gi, oracle_lib, cuda):
    # 156 of the 200 large triangles are clipped, more than one pass of the
    # clip kernel's 16 x 64 threads (7 fan pieces each) covers.  The workspace
    # holds every list without falling back: at most 4 tiles per piece and
    # 7 pieces per clipped triangle fit 4 entries per triangle plus 1024, and
    # 200 x 7 setup records fit 1024 + 8200 / 16.
    mesh, light = clip_heavy_mesh()
    mvp = gi.passes.view_matrices(light)[2]
    n = clipped_count(mesh, mvp)
    assert n * 7 > 16 * 64, n
    light2 = gi.scenes.View((1.0, 9.0, -0.5), (0.0, 0.0, 0.0), 128, 128, fov=75.0)
    mvps = np.stack([mvp, gi.passes.view_matrices(light2)[2]])
    print(f"clip-heavy depth: {mesh.ntris} triangles, {n} clipped in map 0, 2 maps of 128^2", flush=True)
    got, want = _depth_both(gi, oracle_lib, mesh, mvps, 128, 128)
    assert bits_equal(got, want), mismatch_report(got, want, "clip-heavy depth")
    assert (want < 1.0).mean() > 0.5


def test_depth_pass_deterministic(gi, cuda):
    wl = gi.scenes.config3(nsamples=4, sm=2048)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    mvps = light_mvps(wl)
    a = _np(gi.render_shadow_map(dm, mvps, 2048, 2048))
    b = _np(gi.render_shadow_map(dm, mvps, 2048, 2048))
    assert bits_equal(a, b)
    assert a.min() >= 0.0 and a.max() <= 1.0


def test_gbuffer_config1(gi, oracle_lib, cuda):
    wl = gi.scenes.config1()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    mvp = camera_mvp(wl)
    g = gi.render_gbuffer(dm, mvp, 800, 600)
    pos, nrm, _ = oracle_lib.gbuffer(wl.mesh.positions, wl.mesh.normals, wl.mesh.indices, mvp, 800, 600)
    assert bits_equal(_np(g.position), pos), mismatch_report(_np(g.position), pos, "gbuffer position")
    assert bits_equal(_np(g.normal), nrm), mismatch_report(_np(g.normal), nrm, "gbuffer normal")


def test_gbuffer_facing_plane(gi, oracle_lib, cuda):
    """float(gl_FrontFacing) per pixel on a scene seen from both sides of its
    triangles (a soup: random winding)."""
    mesh = shrink_soup(20_000)
    wl = gi.scenes.config2(ntris=10, view_w=640, view_h=360, sm=512)
    wl.mesh = mesh
    dm = gi.DeviceMesh.from_mesh(mesh)
    mvp = camera_mvp(wl)
    g = gi.render_gbuffer(dm, mvp, 640, 360)
    pos, nrm, fac = oracle_lib.gbuffer(mesh.positions, mesh.normals, mesh.indices, mvp, 640, 360)
    assert bits_equal(_np(g.position), pos)
    assert bits_equal(_np(g.normal), nrm)
    assert bits_equal(_np(g.facing), fac), mismatch_report(_np(g.facing), fac, "gbuffer facing")
    hit = pos[0] != 0
    assert 0.01 < fac[hit].mean() < 0.99  # both sides are seen (the ground quad is most of the view)


def _phong_case(gi, oracle_lib, g, shadow, wl, color_plane=None):
    params = gi.PhongParams.from_views(wl.camera, wl.lights[0].eye, 0.25)
    got = _np(gi.shade_scene(g, shadow, params, color=color_plane))
    want = oracle_lib.phong(_np(g.position), _np(g.normal), _np(g.facing), _np(shadow), params.MV,
                            params.normalMatrix, params.lightPosition, 0.25, color=params.color,
                            clear_color=params.clearColor,
                            color_plane=None if color_plane is None else _np(color_plane))
    return got, want


def test_phong_over_hard_shadows_config1(gi, oracle_lib, cuda):
    """shadeScene: deferred Phong over the naive hard-shadow buffer."""
    wl = gi.scenes.config1()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)[:1]
    depth = gi.render_shadow_map(dm, mvps, 1024, 1024)
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    params = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=1024, shadowMapHeight=1024, technique="naive",
                             zbias=wl.zbias, preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos)
    shadow = gi.compute_hard_shadows(g, depth[0], params)
    got, want = _phong_case(gi, oracle_lib, g, shadow, wl)
    assert bits_equal(got, want), mismatch_report(got, want, "phong (hard)")
    bg = _np(g.position)[0] == 0
    assert (got[bg] == np.float32(gi.passes.CLEAR_COLOR)).all()
    assert (got[~bg, 0] > 0).mean() > 0.9


def test_phong_over_soft_shadows_with_color_plane(gi, oracle_lib, cuda):
    """deferredShading: Phong over the area-light soft-shadow buffer of a
    back-and-front-facing soup, with a per-pixel colour plane."""
    mesh = shrink_soup(20_000)
    wl = gi.scenes.config3(nsamples=4, sm=512, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 512, 512)
    soft = gi.render_monte_carlo(g, maps, mvps, gi.SoftParams(pcf=True, kernelOrder=3))
    gen = torch.Generator(device="cpu").manual_seed(11)
    colors = torch.rand((wl.camera.height, wl.camera.width, 4), generator=gen).to("cuda")
    got, want = _phong_case(gi, oracle_lib, g, soft, wl, colors)
    assert bits_equal(got, want), mismatch_report(got, want, "phong (soft, colour plane)")
    # the soup: back faces shade with the raw normal
    dm2 = gi.DeviceMesh.from_mesh(mesh)
    g2 = gi.render_gbuffer(dm2, camera_mvp(wl), wl.camera.width, wl.camera.height)
    sh2 = torch.rand((wl.camera.height, wl.camera.width), generator=gen).to("cuda")
    got, want = _phong_case(gi, oracle_lib, g2, sh2, wl)
    assert bits_equal(got, want), mismatch_report(got, want, "phong (soup)")


def test_phong_rejects(gi, cuda):
    import ctypes

    wl = gi.scenes.config1()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 64, 48)
    shadow = torch.ones((48, 64), device="cuda")
    params = gi.PhongParams.from_views(wl.camera, wl.lights[0].eye)
    out = torch.empty((48 * 64 * 4 + 1,), device="cuda")
    c = g.to_c()
    p = params.to_c()
    L = gi.lib()
    # misaligned output
    assert L.gi_phong_pass(ctypes.byref(c), shadow.data_ptr(), None, ctypes.byref(p), out.data_ptr() + 4,
                           None) == gi._lib.GI_EINVAL
    # no facing plane
    c.facing_dev = None
    assert L.gi_phong_pass(ctypes.byref(c), shadow.data_ptr(), None, ctypes.byref(p), out.data_ptr(),
                           None) == gi._lib.GI_EINVAL
    c = g.to_c()
    p.light_position[0] = float("nan")
    assert L.gi_phong_pass(ctypes.byref(c), shadow.data_ptr(), None, ctypes.byref(p), out.data_ptr(),
                           None) == gi._lib.GI_ENONFINITE


def _hard_case(gi, oracle_lib, wl, technique, pre_eval, smw, smh, extra=None):
    extra = dict(extra or {})
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    cam_mvp = camera_mvp(wl)
    g = gi.render_gbuffer(dm, cam_mvp, wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)[:1]
    depth = gi.render_shadow_map(dm, mvps, smw, smh)
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    params = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=smw, shadowMapHeight=smh, technique=technique,
                             zbias=wl.zbias, preEvaluation=pre_eval, MV=MV, normalMatrix=nm, lightPosition=lpos,
                             **extra)
    got = _np(gi.compute_hard_shadows(g, depth[0], params))
    op = oracle_lib.make_params(gi.biased(mvps[0], wl.zbias), smw, smh, technique,
                                kernel_order=params.kernelOrder, penumbra_size=params.penumbraSize,
                                max_search=params.maxSearch, depth_threshold=params.depthThreshold,
                                shadow_intensity=params.shadowIntensity, pre_evaluation=pre_eval, mv=MV,
                                normal_matrix=nm, light_position=lpos)
    want = oracle_lib.hard_shadow(_np(g.position), _np(g.normal), _np(depth[0]), op)
    return got, want


@pytest.mark.parametrize("technique", ["naive", "pcf"])
@pytest.mark.parametrize("pre_eval", [False, True])
def test_hard_shadow_config1(gi, oracle_lib, cuda, technique, pre_eval):
    wl = gi.scenes.config1()
    got, want = _hard_case(gi, oracle_lib, wl, technique, pre_eval, 1024, 1024,
                           {"kernelOrder": 3, "penumbraSize": 1})
    if technique == "naive":
        assert bits_equal(got, want), mismatch_report(got, want, "naive")
        assert set(np.unique(want).tolist()) <= {0.0, 0.25, 1.0}
    else:
        assert np.abs(got - want).max() <= TOL, mismatch_report(got, want, "pcf")
        assert bits_equal(got, want), mismatch_report(got, want, "pcf")


@pytest.mark.parametrize("technique", ["rbsm", "rpcf", "rbsm_filtered", "rbsm_conservative", "rpcf_conservative"])
def test_rbsm_soup(gi, oracle_lib, cuda, technique):
    mesh = shrink_soup(30_000)
    wl = gi.scenes.config2(ntris=10, view_w=640, view_h=360, sm=1024)
    wl.mesh = mesh
    wl.lights = [gi.scenes.ortho_fit((0.3, -1.0, 0.2), mesh, 1024, 1024)]
    got, want = _hard_case(gi, oracle_lib, wl, technique, False, 1024, 1024,
                           {"maxSearch": 16, "depthThreshold": 2.5e-6, "kernelOrder": 3, "penumbraSize": 1})
    assert np.abs(got - want).max() <= TOL, mismatch_report(got, want, technique)
    assert bits_equal(got, want), mismatch_report(got, want, technique)


@pytest.mark.parametrize("technique", ["rpcf", "rbsm_filtered", "rpcf_conservative"])
@pytest.mark.parametrize("order,penumbra,pre_eval", [(1, 1, False), (2, 3, True), (5, 1, False), (7, 1, True),
                                                     (9, 2, False)])
def test_rpcf_tap_loops(gi, oracle_lib, cuda, technique, order, penumbra, pre_eval):
    """RPCF, FilteredRBSM and conservative RPCF tap loops of other sizes (the
    float loop's tap count: 2, 3, 6, 8 and 10 per axis): one lane per tap,
    16, 7 or 1 pixels per wave group (RPCF and FilteredRBSM pack three
    groups' discontinuity taps for the edge walks; 64 taps fill the wave),
    and 100 taps in passes of 64."""
    mesh = shrink_soup(30_000)
    wl = gi.scenes.config2(ntris=10, view_w=640, view_h=360, sm=1024)
    wl.mesh = mesh
    wl.lights = [gi.scenes.ortho_fit((0.3, -1.0, 0.2), mesh, 1024, 1024)]
    got, want = _hard_case(gi, oracle_lib, wl, technique, pre_eval, 1024, 1024,
                           {"maxSearch": 16, "depthThreshold": 2.5e-6, "kernelOrder": order,
                            "penumbraSize": penumbra})
    assert bits_equal(got, want), mismatch_report(got, want, f"{technique} order {order} penumbra {penumbra}")
    assert ((want > 0.25) & (want < 1.0)).sum() > 100


@pytest.fixture(scope="module")
def config2_frame(gi, oracle_lib, cuda):
    """The RBSM workload at its stated size: 200,000-triangle soup (seed 1)
    over a 40x40 ground quad, directional light with a fitted orthographic
    frustum, 2048^2 map, 1920x1080 view; the depth map is checked against
    the oracle's raster once for every technique below."""
    wl = gi.scenes.config2()
    assert wl.mesh.ntris == 200_002 and (wl.camera.width, wl.camera.height) == (1920, 1080)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 1920, 1080)
    mvps = light_mvps(wl)[:1]
    depth = gi.render_shadow_map(dm, mvps, 2048, 2048)
    want = oracle_lib.depth_raster(wl.mesh.positions, wl.mesh.indices, mvps[0], 2048, 2048)
    assert bits_equal(_np(depth[0]), want), mismatch_report(_np(depth[0]), want, "config-2 depth map")
    return wl, g, mvps, depth


@pytest.mark.parametrize("technique,thr", [("rbsm", 2.5e-6), ("rpcf", 2.5e-6), ("rbsm_filtered", 2.5e-6),
                                           ("rbsm", 5e-4), ("rbsm_filtered", 5e-4)])
def test_rbsm_config2_full_size(gi, oracle_lib, config2_frame, technique, thr):
    """Binary, RPCF and fractional RBSM on the full config-2 frame
    (maxSearch 16, kernelOrder 3, penumbraSize 1) against the oracle over
    every pixel; depthThreshold 2.5e-6 as the workload sets it, and 5e-4, at
    which the walks' depth nudges and the newDepth carry take effect."""
    wl, g, mvps, depth = config2_frame
    print(f"config 2: {technique}, depthThreshold {thr:g}", flush=True)
    params = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=2048, shadowMapHeight=2048, technique=technique,
                             zbias=wl.zbias, maxSearch=16, depthThreshold=thr, kernelOrder=3, penumbraSize=1)
    got = _np(gi.compute_hard_shadows(g, depth[0], params))
    op = oracle_lib.make_params(gi.biased(mvps[0], wl.zbias), 2048, 2048, technique, kernel_order=3,
                                penumbra_size=1, max_search=16, depth_threshold=thr, shadow_intensity=0.25)
    want = oracle_lib.hard_shadow(_np(g.position), _np(g.normal), _np(depth[0]), op)
    fg = _np(g.position)[0] != 0
    assert fg.mean() > 0.5
    assert np.abs(got - want).max() <= TOL, mismatch_report(got, want, technique)
    assert bits_equal(got, want), mismatch_report(got, want, technique)
    # the frame exercises the revectorized branches: shadowed, lit and, for
    # the filtering techniques, fractional pixels are all present
    v = want[fg]
    if technique == "rbsm":
        assert (v == 0.25).mean() > 0.05 and (v == 1.0).mean() > 0.05
    else:  # the soup's shadow is fragmented at texel scale: most 4x4 averages are fractional
        assert (v == 0.25).mean() > 0.05 and (v == 1.0).mean() > 0.005
        assert ((v > 0.25) & (v < 1.0)).mean() > 0.1


def _soft_case(gi, oracle_lib, wl, sm, params):
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    cam_mvp = camera_mvp(wl)
    g = gi.render_gbuffer(dm, cam_mvp, wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, sm, sm)
    got = _np(gi.render_monte_carlo(g, maps, mvps, params, zbias=wl.zbias))
    common, trans = gi.soft_uniforms(mvps, wl.zbias)
    want = oracle_lib.soft_shadow(_np(g.position), _np(maps), common, trans, si=params.shadowIntensity,
                                  pcf=params.pcf, kernel_order=params.kernelOrder, penumbra=params.penumbra,
                                  inclusive=params.inclusive)
    return got, want


@pytest.mark.parametrize("pcf,inclusive,order,penumbra", [
    (False, True, 3, 1.0),   # one tap per sample
    (True, True, 3, 1.0),    # the reference loop: 4x4 taps
    (True, False, 3, 1.0),   # 3x3 taps
    (True, False, 3, 2.0),   # 3 taps spanning more than 2 texels: per-tap path
    (True, True, 7, 1.5),    # 8 taps per axis: generic loop
])
def test_soft_shadow_area_light(gi, oracle_lib, cuda, pcf, inclusive, order, penumbra):
    wl = gi.scenes.config3(nsamples=9, view_w=480, view_h=270, sm=1024, freq=40)
    params = gi.SoftParams(pcf=pcf, inclusive=inclusive, kernelOrder=order, penumbra=penumbra)
    got, want = _soft_case(gi, oracle_lib, wl, 1024, params)
    assert np.abs(got - want).max() <= TOL, mismatch_report(got, want, "soft")
    assert bits_equal(got, want), mismatch_report(got, want, "soft")
    vis = want[want != 0]
    assert vis.min() >= 0.25 - 1e-7 and vis.max() <= 1.0 + 1e-7


@pytest.mark.parametrize("si", [0.3, 0.0, 1.0, 0.5 + 2.0 ** -13, 0.375])
def test_soft_shadow_intensities(gi, oracle_lib, cuda, si):
    """Shadow intensities other than 0.25, including ones whose tap sums
    round in FP32 (0.3, 0.5 + 2^-13): the kernel adds the taps in the loop's
    order, bit-identical to the oracle's sequential sum."""
    wl = gi.scenes.config3(nsamples=9, view_w=480, view_h=270, sm=1024, freq=40)
    for inclusive in (True, False):
        params = gi.SoftParams(shadowIntensity=si, pcf=True, inclusive=inclusive, kernelOrder=3, penumbra=1.0)
        got, want = _soft_case(gi, oracle_lib, wl, 1024, params)
        assert bits_equal(got, want), mismatch_report(got, want, f"soft si={si} inclusive={inclusive}")


def test_soft_shadow_map_border(gi, oracle_lib, cuda):
    """A light frustum narrower than the scene: taps fall off the map edge and
    read the border (depth 0, shadowed)."""
    wl = gi.scenes.config3(nsamples=4, view_w=320, view_h=200, sm=256, freq=20)
    for v in wl.lights:
        v.fov = 20.0
    params = gi.SoftParams(pcf=True, inclusive=False)
    got, want = _soft_case(gi, oracle_lib, wl, 256, params)
    assert bits_equal(got, want), mismatch_report(got, want, "soft border")


def test_soft_shadow_full_size_properties(gi, cuda):
    """1080p, 64 samples, 2048^2 maps: the config-3 frame.  Properties that hold
    at any size: background pixels 0, lit range [si, 1], and the frame equals
    the mean of per-sample single-light frames computed separately."""
    wl = gi.scenes.config3()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 1920, 1080)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 2048, 2048)
    params = gi.SoftParams(pcf=True, inclusive=False)
    full = gi.render_monte_carlo(g, maps, mvps, params)
    bg = g.position[0] == 0
    assert bool((full[bg] == 0).all())
    fg = full[~bg]
    assert float(fg.min()) >= 0.25 - 1e-6 and float(fg.max()) <= 1.0 + 1e-6
    acc = torch.zeros_like(full)
    for i in range(0, 64, 8):
        part = gi.render_monte_carlo(g, maps[i:i + 8].contiguous(), mvps[i:i + 8], params)
        acc += part * 8
    mean = acc / 64
    assert float((mean[~bg] - fg).abs().max()) <= 1e-5


def _rev_frame(gi, nsamples=9):
    """The area-light scene with nsamples corner maps, their discontinuity
    layers (corner eye unrotated in the pre-evaluation, as during the quad-tree
    evaluation) and the uniforms the final pass uses."""
    wl = gi.scenes.config3(nsamples=nsamples, view_w=640, view_h=360, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 640, 360)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 1024, 1024, polygon_offset=(4.0, 20.0))
    biased = np.stack([gi.biased(m) for m in mvps])
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    return wl, g, mvps, maps, biased, MV, nm, lpos


def test_rev_discontinuity_matches_oracle(gi, oracle_lib, cuda):
    """Discontinuity.frag per corner map, and no write past the last layer."""
    wl, g, mvps, maps, biased, MV, nm, _ = _rev_frame(gi, nsamples=4)
    disc = torch.full((3, g.height, g.width), 7, dtype=torch.uint8, device="cuda")
    gp, gn = _np(g.position), _np(g.normal)
    for i in range(4):
        eye = tuple(float(v) for v in wl.lights[i].eye)
        print(f"rev discontinuity map {i}", flush=True)
        gi.passes.rev_discontinuity(g, maps, i, biased[i], MV, nm, eye, 0.25, disc)
        if i < 3:
            op = oracle_lib.make_params(biased[i], 1024, 1024, "naive", shadow_intensity=0.25, max_search=16,
                                        pre_evaluation=True, mv=MV, normal_matrix=nm, light_position=eye)
            want = oracle_lib.rev_discontinuity(gp, gn, _np(maps[i]), op)
            got = disc[i].cpu().numpy()
            assert (got == want).all(), f"map {i}: {int((got != want).sum())} pixels differ"
            assert ((want & 15) != 0).sum() > 100  # the occluder's edges are there
    # slot 3 has no layer: layers 0-2 are as written above, nothing else changed
    assert int((disc == 7).sum()) == 0


def test_rev_node_test_matches_oracle(gi, oracle_lib, cuda):
    """RevectorizationBasedReprojection.frag + Evaluation.frag at level 0 and
    then level 2 (div 2) over the same visibility map, with the layer clamp
    (a 3-layer array, corner slots up to 3)."""
    wl, g, mvps, maps, biased, MV, nm, _ = _rev_frame(gi, nsamples=4)
    gp, gn = _np(g.position), _np(g.normal)
    disc = gi.passes.rev_discontinuity_layers(g, layers=3)
    disc_np = np.zeros((3, g.height, g.width), np.uint8)
    for i in range(3):
        eye = tuple(float(v) for v in wl.lights[i].eye)
        gi.passes.rev_discontinuity(g, maps, i, biased[i], MV, nm, eye, 0.25, disc)
        op = oracle_lib.make_params(biased[i], 1024, 1024, "naive", shadow_intensity=0.25, max_search=16,
                                    pre_evaluation=True, mv=MV, normal_matrix=nm, light_position=eye)
        disc_np[i] = oracle_lib.rev_discontinuity(gp, gn, _np(maps[i]), op)
    vis = torch.zeros((g.height, g.width), dtype=torch.float32, device="cuda")
    vis_np = np.zeros((g.height, g.width), np.float32)
    for level, order, div in ((0, [0, 1, 2, 3], 1), (2, [3, 2, 1, 0], 2), (3, [1, 3, 0, 2], 3),
                              (4, [2, 0, 3, 1], 6)):
        print(f"rev node test level {level}", flush=True)
        img, count = gi.passes.rev_quadtree_node_test(g, maps, order, biased[order], disc, level, vis, div)
        want_img, want_count = oracle_lib.rev_node_test(gp, _np(maps), order, biased[order], disc_np, level, vis_np,
                                                        div)
        assert bits_equal(_np(img), want_img)
        assert bits_equal(_np(vis), vis_np)
        assert count == want_count
        assert (want_img == 5.0).sum() > 0 and (vis_np == 0.5).sum() > 0


@pytest.mark.parametrize("rpcf,si,nsamples", [(False, 0.25, 9), (True, 0.25, 9), (False, 0.5, 9),
                                               (False, 0.25, 81), (True, 0.25, 81), (True, 0.25, 4),
                                               (True, 0.5, 16), (True, 0.25, 36)])
def test_rev_soft_matches_oracle(gi, oracle_lib, cuda, rpcf, si, nsamples):
    """RevectorizationBasedAccurateSoftShadow.frag with mixed quad-tree levels,
    discontinuity layers for the first 6 maps (the rest read a clamped or
    stale layer as the reference does), and the visibility classification of
    one level-0 node test.  Binary: 9 samples put 7 pixels in a wave's lanes,
    81 run one pixel in two passes of 64 and 17 samples.  RPCF (28
    pixel-sample pairs per workgroup pass): 4 samples give 7 pixels per pass,
    9 three, 16 one pixel with 12 pairs idle, 36 one pixel in passes of 28
    and 8, 81 one pixel in passes of 28, 28 and 25.  The node test takes the
    light's four corner samples."""
    wl, g, mvps, maps, biased, MV, nm, lpos = _rev_frame(gi, nsamples=nsamples)
    gp, gn = _np(g.position), _np(g.normal)
    layers = min(6, nsamples)
    k = int(round(nsamples ** 0.5))
    corners = [0, k - 1, k * (k - 1), k * k - 1]
    disc = gi.passes.rev_discontinuity_layers(g, layers=layers)
    disc_np = np.zeros((layers, g.height, g.width), np.uint8)
    for i in range(layers):
        eye = tuple(float(v) for v in wl.lights[i].eye)
        gi.passes.rev_discontinuity(g, maps, i, biased[i], MV, nm, eye, 0.25, disc)
        op = oracle_lib.make_params(biased[i], 1024, 1024, "naive", shadow_intensity=0.25, max_search=16,
                                    pre_evaluation=True, mv=MV, normal_matrix=nm, light_position=eye)
        disc_np[i] = oracle_lib.rev_discontinuity(gp, gn, _np(maps[i]), op)
    vis = torch.zeros((g.height, g.width), dtype=torch.float32, device="cuda")
    vis_np = np.zeros((g.height, g.width), np.float32)
    gi.passes.rev_quadtree_node_test(g, maps, corners, biased[corners], disc, 0, vis)
    oracle_lib.rev_node_test(gp, _np(maps), corners, biased[corners], disc_np, 0, vis_np)
    levels = ([0, 1, 2, 3, 4, 1, 6, 2, 0] * 9)[:nsamples]
    common, trans = gi.passes.adaptive_trans(mvps, levels)
    params = gi.RevSoftParams(MV, nm, lpos, shadowIntensity=si, RPCF=rpcf, maxSearch=16, depthThreshold=2.5e-6)
    print(f"rev soft rpcf={rpcf} si={si}", flush=True)
    got = _np(gi.passes.render_rev_adaptive(g, maps, common, trans, disc, vis, params))
    op = oracle_lib.make_params(common, 1024, 1024, "naive", shadow_intensity=si, max_search=16,
                                depth_threshold=2.5e-6, pre_evaluation=True, mv=MV, normal_matrix=nm,
                                light_position=lpos)
    want = oracle_lib.rev_soft(gp, gn, _np(maps), common, trans, disc_np, vis_np, op, rpcf=rpcf)
    assert bits_equal(got, want), mismatch_report(got, want, f"rev soft rpcf={rpcf} si={si}")
    pen = (vis_np == 0.5) & (gp[0] != 0)
    assert pen.sum() > 100
    assert ((want > si) & (want < 1.0)).sum() > 100


def test_rev_soft_without_penumbra_pixels(gi, oracle_lib, cuda):
    """A visibility map with no pixel at 0.5 leaves the penumbra list empty:
    every pixel takes its clamped classification, and the sample kernel
    launches with nothing to do."""
    wl, g, mvps, maps, biased, MV, nm, lpos = _rev_frame(gi, nsamples=4)
    gp, gn = _np(g.position), _np(g.normal)
    disc = gi.passes.rev_discontinuity_layers(g, layers=4)
    vis_np = np.where((np.arange(g.width)[None, :] // 7) % 2 == 0, 1.0, 0.0).astype(np.float32)
    vis_np = np.broadcast_to(vis_np, (g.height, g.width)).copy()
    vis = torch.from_numpy(vis_np).cuda()
    common, trans = gi.passes.adaptive_trans(mvps, [0, 1, 2, 3])
    disc_np = _np(disc)
    for rpcf in (False, True):
        params = gi.RevSoftParams(MV, nm, lpos, shadowIntensity=0.25, RPCF=rpcf, maxSearch=16)
        print(f"rev soft, empty list, rpcf={rpcf}", flush=True)
        got = _np(gi.passes.render_rev_adaptive(g, maps, common, trans, disc, vis, params))
        op = oracle_lib.make_params(common, 1024, 1024, "naive", shadow_intensity=0.25, max_search=16,
                                    pre_evaluation=True, mv=MV, normal_matrix=nm, light_position=lpos)
        want = oracle_lib.rev_soft(gp, gn, _np(maps), common, trans, disc_np, vis_np, op, rpcf=rpcf)
        assert bits_equal(got, want)
        fg = gp[0] != 0
        assert set(np.unique(want[fg]).tolist()) <= {0.25, 1.0}


def test_rev_rejects(gi, cuda):
    import ctypes

    wl, g, mvps, maps, biased, MV, nm, lpos = _rev_frame(gi, nsamples=4)
    L = gi.lib()
    disc = gi.passes.rev_discontinuity_layers(g, layers=4)
    vis = torch.zeros((g.height, g.width), dtype=torch.float32, device="cuda")
    out = torch.empty((g.height, g.width), dtype=torch.float32, device="cuda")
    c = g.to_c(with_normals=True)
    F16, F9, F3 = ctypes.c_float * 16, ctypes.c_float * 9, ctypes.c_float * 3
    args = (F16(*biased[0].reshape(-1)), F16(*MV.reshape(-1)), F9(*nm.reshape(-1)), F3(*lpos))
    for idx, layers in ((-1, 4), (4, 4), (0, 0)):
        assert L.gi_rev_discontinuity_pass(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, idx, *args,
                                           ctypes.c_float(0.25), disc.data_ptr(), layers, None) == gi._lib.GI_EINVAL
    nan = F16(*([float("nan")] * 16))
    assert L.gi_rev_discontinuity_pass(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, 0, nan, *args[1:],
                                       ctypes.c_float(0.25), disc.data_ptr(), 4, None) == gi._lib.GI_ENONFINITE
    cnt = torch.zeros((1,), dtype=torch.int32, device="cuda")
    idx4 = (ctypes.c_int * 4)(0, 1, 2, 4)
    m64 = (ctypes.c_float * 64)(*np.stack(biased).reshape(-1)[:64])
    assert L.gi_rev_quadtree_node_test(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, idx4, m64, disc.data_ptr(),
                                       4, 0, 1, out.data_ptr(), vis.data_ptr(), cnt.data_ptr(), None) == \
        gi._lib.GI_EINVAL
    common, trans = gi.passes.adaptive_trans(mvps, [0, 0, 0, 0])
    tr = torch.from_numpy(np.ascontiguousarray(trans)).cuda()
    nws = L.gi_rev_soft_workspace_size(g.width, g.height)
    assert nws == 256 + g.width * g.height * 4
    ws = torch.empty((nws,), dtype=torch.uint8, device="cuda")
    p = gi.RevSoftParams(MV, nm, lpos, maxSearch=65).to_c()
    assert L.gi_rev_adaptive_soft_shadow_pass(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, F16(*common),
                                              tr.data_ptr(), disc.data_ptr(), 4, vis.data_ptr(), ctypes.byref(p),
                                              out.data_ptr(), ws.data_ptr(), nws, None) == gi._lib.GI_EINVAL
    p = gi.RevSoftParams(MV, nm, lpos).to_c()
    assert L.gi_rev_adaptive_soft_shadow_pass(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, F16(*common),
                                              tr.data_ptr(), disc.data_ptr(), 4, vis.data_ptr(), ctypes.byref(p),
                                              out.data_ptr(), ws.data_ptr(), nws - 1, None) == gi._lib.GI_ENOSPACE
    assert L.gi_rev_adaptive_soft_shadow_pass(ctypes.byref(c), maps.data_ptr(), 4, 1024, 1024, F16(*common),
                                              tr.data_ptr(), disc.data_ptr(), 4, vis.data_ptr(), ctypes.byref(p),
                                              out.data_ptr(), None, nws, None) == gi._lib.GI_EINVAL


def _pcss_scene(gi, sm):
    """The area-light scene lit from 40x further away with a 5-degree light
    frustum: window depths near 0.997, above PCSS's 0.99 blocker cut-off."""
    wl = gi.scenes.config3(nsamples=1, view_w=640, view_h=360, sm=sm)
    wl.lights = [gi.scenes.View((120.0, 280.0, 80.0), (0.0, 0.0, 0.0), sm, sm, fov=5.0)]
    return wl


@pytest.mark.parametrize("sm,radius,znear,pre_eval,si,ks", [(1024, 8, 1, True, 0.25, 15), (2048, 8, 1, False, 0.25, 15),
                                                            (2048, 256, 1, True, 0.25, 15), (1024, 64, 0, True, 0.25, 15),
                                                            (2048, 256, 1, True, 0.3, 15), (2048, 256, 1, True, 0.25, 17)])
def test_pcss_matches_oracle(gi, oracle_lib, cuda, sm, radius, znear, pre_eval, si, ks):
    """PlausibleSoftShadow.frag with PCSS == 1, also with a shadow intensity
    off the 2^-16 grid (0.3) and a 17 x 17 PCF kernel."""
    wl = _pcss_scene(gi, sm)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)
    depth = gi.render_shadow_map(dm, mvps, sm, sm, polygon_offset=(4.0, 20.0))
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    params = gi.PCSSParams(lightMVP=mvps[0], shadowMapWidth=sm, shadowMapHeight=sm, lightSourceRadius=radius,
                           zNear=znear, preEvaluation=pre_eval, MV=MV, normalMatrix=nm, lightPosition=lpos,
                           shadowIntensity=si, kernelSize=ks)
    got = _np(gi.render_soft_shadows(g, depth[0], params))
    op = oracle_lib.make_params(gi.biased(mvps[0]), sm, sm, "naive", shadow_intensity=si,
                                pre_evaluation=pre_eval, mv=MV, normal_matrix=nm, light_position=lpos)
    want = oracle_lib.pcss(_np(g.position), _np(g.normal) if pre_eval else None, _np(depth[0]), op,
                           blocker_search_size=7, kernel_size=ks, light_source_radius=radius, z_near=znear)
    assert bits_equal(got, want), mismatch_report(got, want, f"pcss sm={sm} radius={radius} zNear={znear}")
    geo = _np(g.position)[0] != 0
    if znear == 0 and not pre_eval:
        assert (want[geo] == 1.0).all()
    if radius == 256:
        pen = (want > si) & (want < 1.0)
        assert pen.mean() > 0.001  # a penumbra exists


def test_pcss_rejects(gi, cuda):
    import ctypes

    wl = gi.scenes.config1()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 64, 48)
    depth = torch.ones((64, 64), device="cuda")
    out = torch.empty((48, 64), device="cuda")
    L = gi.lib()
    for field, val in (("kernel_size", 1), ("blocker_search_size", 1), ("kernel_size", 65),
                       ("light_source_radius", -1)):
        p = gi.PCSSParams(lightMVP=light_mvps(wl)[0], shadowMapWidth=64, shadowMapHeight=64).to_c()
        setattr(p, field, val)
        c = g.to_c(with_normals=False)
        assert L.gi_pcss_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), None) == \
            gi._lib.GI_EINVAL, field
    p = gi.PCSSParams(lightMVP=light_mvps(wl)[0], shadowMapWidth=64, shadowMapHeight=64).to_c()
    p.light_mvp[3] = float("inf")
    c = g.to_c(with_normals=False)
    assert L.gi_pcss_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), None) == \
        gi._lib.GI_ENONFINITE


# depth thresholds: the `d` values of the reference's scene configurations
@pytest.mark.parametrize("sm,radius,znear,pre_eval,max_search,thr", [(1024, 8, 1, True, 16, 2.5e-6),
                                                                     (2048, 64, 1, False, 16, 2.5e-5),
                                                                     (2048, 256, 1, True, 16, 2.5e-6),
                                                                     (1024, 64, 1, True, 4, 2.5e-5),
                                                                     (1024, 64, 0, True, 16, 0.0)])
def test_rbssm_matches_oracle(gi, oracle_lib, cuda, sm, radius, znear, pre_eval, max_search, thr):
    wl = _pcss_scene(gi, sm)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)
    depth = gi.render_shadow_map(dm, mvps, sm, sm, polygon_offset=(4.0, 20.0))
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    params = gi.RBSSMParams(lightMVP=mvps[0], shadowMapWidth=sm, shadowMapHeight=sm, lightSourceRadius=radius,
                            zNear=znear, preEvaluation=pre_eval, MV=MV, normalMatrix=nm, lightPosition=lpos,
                            maxSearch=max_search, depthThreshold=thr)
    print(f"rbssm sm={sm} radius={radius} zNear={znear} maxSearch={max_search}", flush=True)
    got = _np(gi.render_rbssm(g, depth[0], params))
    op = oracle_lib.make_params(gi.biased(mvps[0]), sm, sm, "naive", shadow_intensity=0.25, max_search=max_search,
                                depth_threshold=thr, pre_evaluation=pre_eval, mv=MV, normal_matrix=nm,
                                light_position=lpos)
    want = oracle_lib.rbssm(_np(g.position), _np(g.normal) if pre_eval else None, _np(depth[0]), op,
                            blocker_search_size=7, kernel_size=7, light_source_radius=radius, z_near=znear)
    assert bits_equal(got, want), mismatch_report(got, want, f"rbssm sm={sm} radius={radius} zNear={znear}")
    geo = _np(g.position)[0] != 0
    if znear == 0:
        lit = geo if not pre_eval else geo & (want != 0.25)
        assert (want[lit] == 1.0).all()
    if radius >= 64 and znear == 1:
        pen = (want > 0.25) & (want < 1.0)
        assert pen.mean() > 0.001  # a penumbra exists
        # revectorized taps differ from PCF over the same taps somewhere
        pcf = oracle_lib.pcss(_np(g.position), _np(g.normal) if pre_eval else None, _np(depth[0]), op,
                              blocker_search_size=7, kernel_size=7, light_source_radius=radius, z_near=znear)
        assert (pcf != want).any()


@pytest.mark.parametrize("ks,smw,smh", [(15, 1024, 1024), (4, 1024, 1024), (7, 1536, 1024), (2, 1024, 1024),
                                        (9, 1024, 1024)])
def test_rbssm_kernel_sizes_and_map_shapes(gi, oracle_lib, cuda, ks, smw, smh):
    """RBSSM with the reference's other kernel sizes (15; an even size runs
    int(-1.5) = -1 .. 1 and divides by 16) and a non-square map.  The tap
    lanes: 1 tap (size 2) puts 64 pixels in a wave, 9 taps 7 pixels, 49 one
    pixel; 81 and 225 taps run one pixel in passes of 64."""
    wl = gi.scenes.config3(nsamples=1, view_w=480, view_h=270, sm=smw, freq=71)
    wl.lights = [gi.scenes.View((120.0, 280.0, 80.0), (0.0, 0.0, 0.0), smw, smh, fov=5.0)]
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), wl.camera.width, wl.camera.height)
    mvps = light_mvps(wl)
    depth = gi.render_shadow_map(dm, mvps, smw, smh, polygon_offset=(4.0, 20.0))
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    params = gi.RBSSMParams(lightMVP=mvps[0], shadowMapWidth=smw, shadowMapHeight=smh, lightSourceRadius=64,
                            kernelSize=ks, preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos,
                            maxSearch=16, depthThreshold=2.5e-5)
    print(f"rbssm ks={ks} map {smw}x{smh}", flush=True)
    got = _np(gi.render_rbssm(g, depth[0], params))
    op = oracle_lib.make_params(gi.biased(mvps[0]), smw, smh, "naive", shadow_intensity=0.25, max_search=16,
                                depth_threshold=2.5e-5, pre_evaluation=True, mv=MV, normal_matrix=nm,
                                light_position=lpos)
    want = oracle_lib.rbssm(_np(g.position), _np(g.normal), _np(depth[0]), op, blocker_search_size=7,
                            kernel_size=ks, light_source_radius=64, z_near=1)
    assert bits_equal(got, want), mismatch_report(got, want, f"rbssm ks={ks} {smw}x{smh}")
    if ks > 2:
        assert ((want > 0.25) & (want < 1.0)).sum() > 100
    else:  # one tap divided by 4: the loop's pixels end in [si / 4, 1 / 4]
        assert ((want > 0.0) & (want < 0.25)).sum() > 100


def test_rbssm_rejects(gi, cuda):
    import ctypes

    wl = gi.scenes.config1()
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 64, 48)
    depth = torch.ones((64, 64), device="cuda")
    out = torch.empty((48, 64), device="cuda")
    L = gi.lib()
    nws = L.gi_rbssm_workspace_size(64, 48)
    assert nws == 256 + 64 * 48 * 8
    assert L.gi_rbssm_workspace_size(0, 48) == 0
    ws = torch.empty((nws,), dtype=torch.uint8, device="cuda")
    base = gi.RBSSMParams(lightMVP=light_mvps(wl)[0], shadowMapWidth=64, shadowMapHeight=64)
    for field, val in (("kernel_size", 1), ("kernel_size", 32), ("blocker_search_size", 1),
                       ("light_source_radius", -1), ("max_search", -1), ("max_search", 65)):
        p = base.to_c()
        if field == "max_search":
            p.max_search = val
        else:
            setattr(p.pcss, field, val)
        c = g.to_c(with_normals=False)
        assert L.gi_rbssm_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), ws.data_ptr(),
                               nws, None) == gi._lib.GI_EINVAL, (field, val)
    p = base.to_c()
    p.depth_threshold = float("nan")
    c = g.to_c(with_normals=False)
    assert L.gi_rbssm_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), ws.data_ptr(), nws,
                           None) == gi._lib.GI_ENONFINITE
    p = base.to_c()
    assert L.gi_rbssm_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), ws.data_ptr(),
                           nws - 1, None) == gi._lib.GI_ENOSPACE
    assert L.gi_rbssm_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(), None, nws,
                           None) == gi._lib.GI_EINVAL
    assert L.gi_rbssm_pass(ctypes.byref(c), depth.data_ptr(), ctypes.byref(p), out.data_ptr(),
                           ws.data_ptr() + 4, nws - 4, None) == gi._lib.GI_EINVAL


def _write_obj(path, mesh):
    lines = ["v " + " ".join(repr(float(c)) for c in p) for p in mesh.positions]
    lines += ["f " + " ".join(str(int(i) + 1) for i in t) for t in mesh.indices]
    path.write_text("\n".join(lines) + "\n")


def test_loaded_scene_through_the_pipeline(gi, oracle_lib, cuda, tmp_path):
    """A scene configuration in the reference's format (OBJ files, s/t/r/c
    lines, camera, light and depth threshold) loaded by libgi's loader, then
    the ShadowMapping frame on it: depth map with glPolygonOffset(4, 20),
    G-buffer, hard shadows (naive and conservative RBSM) and Phong, each
    bit-identical to the oracle on the loaded mesh."""
    sphere = gi.scenes.displaced_icosphere(12, seed=7)
    _write_obj(tmp_path / "sphere.obj", sphere)
    quad = gi.scenes.ground_quad(1.0, 1.0, 0.0)
    _write_obj(tmp_path / "plane.obj", quad)
    (tmp_path / "scene.txt").write_text("""o sphere.obj
s 4.0 4.0 4.0
r 0.0 30.0 0.0
t 0.0 4.0 10.0
c 0.75 0.6 0.0
+
o plane.obj
s 60.0 1.0 40.0
t 0.0 -8.0 0.0
+
ve 0.0 41.0 -50.0
va 0.0 16.0 -10.0
le 10.0 130.0 100.0
la 0.0 0.0 0.0
d 0.000025
""")
    wl = gi.scenes.load_scene(tmp_path / "scene.txt", base_dir=tmp_path, window=(640, 480), sm=1024)
    assert wl.mesh.ntris == sphere.ntris + quad.ntris
    print(f"loaded scene: {wl.mesh.nverts} vertices, {wl.mesh.ntris} triangles", flush=True)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    mvps = light_mvps(wl)
    depth = gi.render_shadow_map(dm, mvps, 1024, 1024, polygon_offset=(4.0, 20.0))
    want_depth = oracle_lib.depth_raster(wl.mesh.positions, wl.mesh.indices, mvps[0], 1024, 1024, (4.0, 20.0))
    assert bits_equal(_np(depth[0]), want_depth)
    cam = camera_mvp(wl)
    g = gi.render_gbuffer(dm, cam, 640, 480)
    pos, nrm, fac = oracle_lib.gbuffer(wl.mesh.positions, wl.mesh.normals, wl.mesh.indices, cam, 640, 480)
    assert bits_equal(_np(g.position), pos) and bits_equal(_np(g.normal), nrm) and bits_equal(_np(g.facing), fac)
    MV, nm, lpos = gi.phong_uniforms(wl.camera, wl.lights[0].eye)
    for tech in ("naive", "rbsm_conservative"):
        params = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=1024, shadowMapHeight=1024, technique=tech,
                                 depthThreshold=wl.extra["depth_threshold"], preEvaluation=True, MV=MV,
                                 normalMatrix=nm, lightPosition=lpos)
        got = gi.compute_hard_shadows(g, depth[0], params)
        op = oracle_lib.make_params(gi.biased(mvps[0]), 1024, 1024, tech, depth_threshold=params.depthThreshold,
                                    shadow_intensity=0.25, pre_evaluation=True, mv=MV, normal_matrix=nm,
                                    light_position=lpos)
        want = oracle_lib.hard_shadow(pos, nrm, want_depth, op)
        assert bits_equal(_np(got), want), mismatch_report(_np(got), want, f"loaded scene {tech}")
    assert (want[pos[0] != 0] == 0.25).mean() > 0.01  # the sphere shadows the plane
    rgba = gi.shade_scene(g, got, gi.PhongParams.from_views(wl.camera, wl.lights[0].eye))
    want_rgba = oracle_lib.phong(pos, nrm, fac, want, MV, nm, lpos, 0.25)
    assert bits_equal(_np(rgba), want_rgba)


@pytest.mark.parametrize("div", [1, 4])
def test_quadtree_node_test_matches_oracle(gi, oracle_lib, cuda, div):
    """One quad-tree node's test on the area-light scene: the four corner
    maps of a 0.5-wide light, the reprojection image and the pass count."""
    wl = gi.scenes.config3(nsamples=4, view_w=640, view_h=360, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 640, 360)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 1024, 1024, polygon_offset=(4.0, 20.0))
    biased = np.stack([gi.biased(m) for m in mvps])
    img, count = gi.passes.quadtree_node_test(g, maps, [0, 1, 2, 3], biased, div)
    want_img, want_count = oracle_lib.qt_node_test(_np(g.position), _np(maps), [0, 1, 2, 3], biased, div)
    assert bits_equal(_np(img), want_img)
    assert count == want_count and count > 0


@pytest.mark.parametrize("pcf", [False, True])
def test_adaptive_soft_matches_oracle(gi, oracle_lib, cuda, pcf):
    """The weighted AccurateSoftShadow pass with mixed quad-tree levels
    (0-4 and one clamped level 6)."""
    wl = gi.scenes.config3(nsamples=9, view_w=640, view_h=360, sm=1024, freq=40)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, camera_mvp(wl), 640, 360)
    mvps = light_mvps(wl)
    maps = gi.render_shadow_map(dm, mvps, 1024, 1024)
    levels = [0, 1, 2, 3, 4, 1, 6, 2, 0]
    common, trans = gi.passes.adaptive_trans(mvps, levels)
    params = gi.SoftParams(pcf=pcf, kernelOrder=3, penumbra=1.0, inclusive=True)
    got = _np(gi.passes.render_adaptive(g, maps, common, trans, params))
    want = oracle_lib.soft_shadow_adaptive(_np(g.position), _np(maps), common, trans, si=0.25, pcf=pcf,
                                           kernel_order=3, penumbra=1.0, inclusive=True)
    assert bits_equal(got, want), mismatch_report(got, want, f"adaptive soft pcf={pcf}")
