"""The C++ host pass interface (globalillumination_amd/host/shadow_passes.*),
driven like the reference apps' display(): every technique of the hard-shadow
frame and a 9-sample renderMonteCarlo frame, checked against the oracle on the
inputs the C++ side dumped (mesh, matrices, maps, G-buffer)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "host_pipeline_test"
W, H, SM = 320, 240, 512


def load(d, name, dtype=np.float32):
    return np.fromfile(d / name, dtype=dtype)


@pytest.fixture(scope="module")
def dumps(tmp_path_factory, cuda):
    out = tmp_path_factory.mktemp("host")
    assert BIN.exists(), "build() compiles tests/cpp/host_pipeline_test"
    print(f"running {BIN.name}", flush=True)
    res = subprocess.run([str(BIN), str(out)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return out


def test_host_depth_and_gbuffer(dumps, oracle_lib):
    pos, nrm = load(dumps, "mesh_pos.f32"), load(dumps, "mesh_nrm.f32")
    idx = load(dumps, "mesh_idx.i32", np.int32)
    light = load(dumps, "light_mvp.f32")
    depth = load(dumps, "depth.f32").reshape(SM, SM)
    want = oracle_lib.depth_raster(pos, idx, light, SM, SM, (4.0, 20.0))  # glPolygonOffset(4, 20)
    assert (depth.view(np.uint32) == want.view(np.uint32)).all()
    cam = load(dumps, "camera_mvp.f32")
    gp, gn, _ = oracle_lib.gbuffer(pos, nrm, idx, cam, W, H)
    assert (load(dumps, "gbuf_pos.f32").reshape(3, H, W).view(np.uint32) == gp.view(np.uint32)).all()
    assert (load(dumps, "gbuf_nrm.f32").reshape(3, H, W).view(np.uint32) == gn.view(np.uint32)).all()


@pytest.mark.parametrize("technique", ["naive", "pcf", "rbsm", "rpcf", "rbsm_filtered", "rbsm_conservative",
                                       "rpcf_conservative"])
def test_host_compute_hard_shadows(dumps, oracle_lib, gi, technique):
    from globalillumination_amd import camera

    light = load(dumps, "light_mvp.f32")
    depth = load(dumps, "depth.f32").reshape(SM, SM)
    gpos = load(dumps, "gbuf_pos.f32").reshape(3, H, W)
    gnrm = load(dumps, "gbuf_nrm.f32").reshape(3, H, W)
    mv = load(dumps, "camera_mv.f32")
    nm = camera.normal_matrix(mv)
    p = oracle_lib.make_params(camera.mat4_mul(camera.bias_matrix(0.005), light), SM, SM, technique,
                               kernel_order=3, penumbra_size=1, max_search=16, depth_threshold=2.5e-6,
                               shadow_intensity=0.25, pre_evaluation=True, mv=mv, normal_matrix=nm,
                               light_position=(-4.0, 8.0, -4.0))
    want = oracle_lib.hard_shadow(gpos, gnrm, depth, p)
    got = load(dumps, f"hard_{technique}.f32").reshape(H, W)
    assert np.abs(got - want).max() <= 1e-5
    assert (got.view(np.uint32) == want.view(np.uint32)).all()
    assert (got != 0).mean() > 0.5  # most of the view is geometry
    # shadeScene: deferred Phong over this technique's shadow buffer
    pos, nrm = load(dumps, "mesh_pos.f32"), load(dumps, "mesh_nrm.f32")
    _, _, fac = oracle_lib.gbuffer(pos, nrm, load(dumps, "mesh_idx.i32", np.int32), load(dumps, "camera_mvp.f32"),
                                   W, H)
    # the normal matrix the C++ side uploaded (an FP64 inverse rounded once)
    nm_host = load(dumps, "camera_nm.f32")
    assert np.abs(nm_host - nm).max() <= 1e-6
    want_rgba = oracle_lib.phong(gpos, gnrm, fac, want, mv, nm_host, (-4.0, 8.0, -4.0), 0.25)
    got_rgba = load(dumps, f"phong_{technique}.f32").reshape(H, W, 4)
    assert (got_rgba.view(np.uint32) == want_rgba.view(np.uint32)).all()


def test_host_render_monte_carlo(dumps, oracle_lib, gi):
    from globalillumination_amd import camera

    mvps = load(dumps, "soft_light_mvps.f32").reshape(9, 16)
    # UniformSampledLightSource: eye and look-at shifted by the same grid offset
    for i in range(9):
        eye = oracle_lib.uniform_sampling((4.0, 8.0, 4.0), 1, 9, i)
        at = oracle_lib.uniform_sampling((0.0, 0.0, 0.0), 1, 9, i)
        want = oracle_lib.mat4_mul(oracle_lib.perspective(45.0, 1.0, 1.0, 1000.0),
                                   oracle_lib.lookat(eye, at, (0.0, 0.0, 1.0)))
        assert (mvps[i].view(np.uint32) == want.view(np.uint32)).all()
    maps = load(dumps, "soft_maps.f32").reshape(9, SM, SM)
    pos, idx = load(dumps, "mesh_pos.f32"), load(dumps, "mesh_idx.i32", np.int32)
    for i in (0, 8):
        want = oracle_lib.depth_raster(pos, idx, mvps[i], SM, SM, (4.0, 20.0))
        assert (maps[i].view(np.uint32) == want.view(np.uint32)).all()
    biased = np.stack([camera.mat4_mul(camera.bias_matrix(0.0), m) for m in mvps])
    trans = biased[:, 12:16]
    gpos = oracle_lib.gbuffer(pos, load(dumps, "mesh_nrm.f32"), idx,
                              oracle_lib.mat4_mul(oracle_lib.perspective(60.0, W / H, 1.0, 1000.0),
                                                  oracle_lib.lookat((0, 3, 6), (0, 0, 0), (0, 0, 1))), W, H)[0]
    want = oracle_lib.soft_shadow(gpos, maps, biased[-1], trans, si=0.25, pcf=True, kernel_order=3, penumbra=1.0,
                                  inclusive=True)
    got = load(dumps, "soft.f32").reshape(H, W)
    assert (got.view(np.uint32) == want.view(np.uint32)).all()
    # deferredShading over the soft buffer
    from globalillumination_amd import camera

    gp, gn, fac = oracle_lib.gbuffer(pos, load(dumps, "mesh_nrm.f32"), idx,
                                     oracle_lib.mat4_mul(oracle_lib.perspective(60.0, W / H, 1.0, 1000.0),
                                                         oracle_lib.lookat((0, 3, 6), (0, 0, 0), (0, 0, 1))), W, H)
    mv = load(dumps, "soft_camera_mv.f32")
    # V * M with M = I: equal values (signed zeros of V may become +0)
    assert (mv == oracle_lib.lookat((0, 3, 6), (0, 0, 0), (0, 0, 1))).all()
    nm = load(dumps, "soft_camera_nm.f32")
    assert np.abs(nm - camera.normal_matrix(mv)).max() <= 1e-6
    assert (gp.view(np.uint32) == gpos.view(np.uint32)).all()
    want_rgba = oracle_lib.phong(gp, gn, fac, want, mv, nm, (-4.0, 8.0, -4.0), 0.25)
    got_rgba = load(dumps, "soft_phong.f32").reshape(H, W, 4)
    assert (got_rgba.view(np.uint32) == want_rgba.view(np.uint32)).all()


def test_host_render_soft_shadows_pcss(dumps, oracle_lib, gi):
    """SoftShadowPipeline::renderSoftShadows: light-centre depth map with
    glPolygonOffset(4, 20), then PCSS (blocker search 7, kernel 15, radius
    16/2, zNear 1) with the pre-evaluation."""
    from globalillumination_amd import camera

    pos, nrm = load(dumps, "mesh_pos.f32"), load(dumps, "mesh_nrm.f32")
    idx = load(dumps, "mesh_idx.i32", np.int32)
    light = load(dumps, "pcss_light_mvp.f32")
    want_light = oracle_lib.mat4_mul(oracle_lib.perspective(45.0, 1.0, 1.0, 1000.0),
                                     oracle_lib.lookat((60.0, 120.0, 60.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)))
    assert (light.view(np.uint32) == want_light.view(np.uint32)).all()
    depth = load(dumps, "pcss_map.f32").reshape(SM, SM)
    want_depth = oracle_lib.depth_raster(pos, idx, light, SM, SM, (4.0, 20.0))
    assert (depth.view(np.uint32) == want_depth.view(np.uint32)).all()
    cam = oracle_lib.mat4_mul(oracle_lib.perspective(60.0, W / H, 1.0, 1000.0),
                              oracle_lib.lookat((0, 3, 6), (0, 0, 0), (0, 0, 1)))
    gp, gn, _ = oracle_lib.gbuffer(pos, nrm, idx, cam, W, H)
    mv = load(dumps, "soft_camera_mv.f32")
    nm = load(dumps, "soft_camera_nm.f32")
    p = oracle_lib.make_params(camera.mat4_mul(camera.bias_matrix(0.0), light), SM, SM, "naive",
                               shadow_intensity=0.25, pre_evaluation=True, mv=mv, normal_matrix=nm,
                               light_position=(-60.0, 120.0, -60.0))
    want = oracle_lib.pcss(gp, gn, depth, p, blocker_search_size=7, kernel_size=15, light_source_radius=8, z_near=1)
    got = load(dumps, "pcss.f32").reshape(H, W)
    assert (got.view(np.uint32) == want.view(np.uint32)).all()
    geo = gp[0] != 0
    assert (want[geo] < 1.0).mean() > 0.01  # the sphere's shadow is there
