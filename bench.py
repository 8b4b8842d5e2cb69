"""Shadow-pipeline benchmark: shaded Mpix/s of the 1080p, 64-sample area-light
frame (64 light-view depth maps of 2048^2 + the soft-shadow screen pass with
3x3 PCF per sample + deferred Phong shading) on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg3]

Prints one JSON line.  A step is one whole frame (renderMonteCarlo then
deferredShading): every depth map is rasterized, the soft-shadow screen pass
runs and the frame is Phong-shaded, with the mesh, G-buffer and light
matrices already resident in HBM.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = ROOT / "profiles" / "hbm_traffic_area_light_64x2048.md"


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` as recorded by the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE runs, scaled by the
    copy-kernel calibration); None when the summary has no such line."""
    import re

    if not TRAFFIC_SUMMARY.exists():
        return None
    m = re.search(rf"^traffic_bytes {re.escape(kernel)} (\d+)\s*$", TRAFFIC_SUMMARY.read_text(), re.M)
    return int(m.group(1)) if m else None


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg3", choices=["cfg1", "cfg2", "cfg3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=64, help="light samples in the CPU baseline sample")
    ap.add_argument("--no-footprint", action="store_true", help="skip counting the unique texels gathered")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-pass timing")
    ap.add_argument("--diff-steps", type=int, default=10, help="timed steps of the differentiable pass")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-frames timing")
    ap.add_argument("--batch-steps", type=int, default=5, help="timed steps of the batched frames")
    ap.add_argument("--no-passes", action="store_true", help="skip the other screen passes' timing")
    ap.add_argument("--no-rbsm", action="store_true", help="skip the RBSM workload's timing")
    ap.add_argument("--passes-steps", type=int, default=5,
                    help="timed steps per run of each other screen pass (median of 3 runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (scale for the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def touched_texels(gi, torch, g, mvps, smw, smh, tap_offsets, zbias=0.0):
    """Unique shadow-map texels the screen pass reads, per map: the
    algorithmic footprint of the gather, computed independently of the kernel
    with elementwise torch ops."""
    common, trans = gi.soft_uniforms(mvps, zbias)
    dev = g.position.device
    pos = g.position.reshape(3, -1)
    fg = pos[0] != 0
    x, y, z = pos[0][fg], pos[1][fg], pos[2][fg]
    m = torch.tensor(common, device=dev)
    cm = [(m[0 * 4 + i] * x + m[1 * 4 + i] * y) + m[2 * 4 + i] * z for i in range(4)]
    total = 0
    offs = torch.tensor(tap_offsets, dtype=torch.float32, device=dev)
    incr_w, incr_h = 1.0 / smw, 1.0 / smh
    for s in range(trans.shape[0]):
        T = trans[s]
        cw = cm[3] + float(T[3])
        sx = (cm[0] + float(T[0])) / cw
        sy = (cm[1] + float(T[1])) / cw
        mark = torch.zeros(smh * smw, dtype=torch.bool, device=dev)
        for ow in offs:
            u = torch.floor((sx + ow * incr_w) * smw)
            for oh in offs:
                v = torch.floor((sy + oh * incr_h) * smh)
                ok = (u >= 0) & (u < smw) & (v >= 0) & (v < smh)
                mark[(v[ok] * smw + u[ok]).long()] = True
        total += int(mark.sum())
    return total


def tap_offsets(kernel_order, penumbra, inclusive):
    import numpy as np

    off = np.float32(penumbra)
    step = np.float32(np.float32(2.0) * off / np.float32(kernel_order))
    out, w = [], -off
    while (w <= off) if inclusive else (w < off):
        out.append(float(w))
        w = np.float32(w + step)
    return out


def hard_footprint(torch, g, biased, smw, smh, taps, neighbours=False):
    """Unique shadow-map texels a hard-shadow pass reads at its taps
    (per-axis offsets `taps` in texels, as the shader's loop forms them) and,
    for the RBSM family, at the four neighbours of each tap that
    computeDiscontinuity reads.  The edge walks beyond those are not counted,
    so for RBSM the figure is a lower bound of the pass's bytes."""
    dev = g.position.device
    pos = g.position.reshape(3, -1)
    fg = pos[0] != 0
    x, y, z = pos[0][fg], pos[1][fg], pos[2][fg]
    m = torch.tensor(list(biased), dtype=torch.float32, device=dev)
    c = [((m[0 * 4 + i] * x + m[1 * 4 + i] * y) + m[2 * 4 + i] * z) + m[3 * 4 + i] for i in range(4)]
    sx, sy = c[0] / c[3], c[1] / c[3]
    mark = torch.zeros(smh * smw, dtype=torch.bool, device=dev)
    nb = [(0, 0)] + ([(1, 0), (-1, 0), (0, 1), (0, -1)] if neighbours else [])
    for a in taps:
        u = torch.floor((sx + float(a) / smw) * smw)
        for b in taps:
            v = torch.floor((sy + float(b) / smh) * smh)
            for du, dv in nb:
                uu, vv = u + du, v + dv
                ok = (uu >= 0) & (uu < smw) & (vv >= 0) & (vv < smh)
                mark[(vv[ok] * smw + uu[ok]).long()] = True
    return int(mark.sum())


def event_ms(torch, fn, steps, stream, reps=3, warm=3):
    """Milliseconds per call of fn on `stream`: `warm` untimed calls, then the
    median over `reps` runs of `steps` calls bracketed by HIP events."""
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    runs = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        runs.append(e0.elapsed_time(e1) / steps)
    return sorted(runs)[len(runs) // 2]


def config2_timing(gi, torch, steps):
    """RBSM hard shadows on the RBSM workload at its stated size: a
    200,000-triangle soup (seed 1) over a 40x40 quad, directional light with a
    fitted orthographic frustum, 2048^2 FP32 map, 1920x1080 view, maxSearch
    16, depthThreshold 2.5e-6, pre-evaluation on (NonConservativeSMSR.frag
    and FilteredRBSM.frag call it).  Binary RBSM, binary RPCF and fractional
    RBSM (FilteredRBSM's 4 x 4 loop), each timed with HIP events on the
    current stream over `steps` launches; inputs resident.  Algorithmic bytes
    per launch: G-buffer position + normal (24 B) and the R32F output (4 B)
    per foreground pixel, 8 B per background pixel (position x read, 0
    written), and 4 B per unique texel of hard_footprint."""
    import numpy as np

    from globalillumination_amd import passes

    wl = gi.scenes.config2()
    W, H = wl.camera.width, wl.camera.height
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, passes.view_matrices(wl.camera)[2], W, H)
    mvps = np.stack([passes.view_matrices(v)[2] for v in wl.lights])
    depth = passes.render_shadow_map(dm, mvps, 2048, 2048)
    MV, nm, lpos = passes.phong_uniforms(wl.camera, wl.lights[0].eye)
    out = torch.empty((H, W), dtype=torch.float32, device=depth.device)
    stream = torch.cuda.current_stream()
    nfg = int((g.position[0] != 0).sum())
    gbytes = nfg * (24 + 4) + (W * H - nfg) * 8
    biased = passes.biased(mvps[0], wl.zbias)
    res = {"workload": f"RBSM hard shadows: {wl.mesh.ntris:,}-triangle soup + 40x40 ground, 2048^2 map, "
                       f"{W}x{H} view, ortho directional light, maxSearch 16", "steps": steps,
           "foreground_pixels": nfg}
    for tech, taps in (("rbsm", [0.0]), ("rpcf", tap_offsets(3, 1.0, True)), ("rbsm_filtered", tap_offsets(3, 1.0, True))):
        sp = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=2048, shadowMapHeight=2048, technique=tech,
                             zbias=wl.zbias, preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos,
                             maxSearch=16, depthThreshold=2.5e-6, kernelOrder=3, penumbraSize=1)
        ms = event_ms(torch, lambda: passes.compute_hard_shadows(g, depth[0], sp, out=out), steps, stream)
        texels = hard_footprint(torch, g, biased, 2048, 2048, taps, neighbours=True)
        nbytes = gbytes + texels * 4
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[tech] = {"ms": round(ms, 4), "Mpix_per_s": round(W * H / (ms * 1e-3) / 1e6, 1),
                     "bytes": int(nbytes), "unique_texels": texels, "GB/s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return res


def cpu_baseline(wl, mvps, g_pos, g_nrm, g_fac, phong, smw, samples, gi):
    """The oracle (FP32 C port of the reference passes, OpenMP) on a bounded
    sample of the frame: `samples` light samples, full 1080p view, plus the
    Phong pass once (it runs once per frame)."""
    import numpy as np

    import oracle

    # every core this process may run on (the affinity mask), as OpenMP threads
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    oracle.set_threads(cores)
    sub = mvps[:samples]
    t0 = time.perf_counter()
    maps = np.stack([oracle.depth_raster(wl.mesh.positions, wl.mesh.indices, m, smw, smw) for m in sub])
    t_depth = time.perf_counter() - t0
    common, trans = gi.soft_uniforms(sub, wl.zbias)
    t0 = time.perf_counter()
    oracle.soft_shadow(g_pos, maps, common, trans, si=0.25, pcf=True, kernel_order=3, penumbra=1.0,
                       inclusive=False)
    t_screen = time.perf_counter() - t0
    shadow = np.ones(g_pos.shape[1:], np.float32)
    t0 = time.perf_counter()
    oracle.phong(g_pos, g_nrm, g_fac, shadow, phong.MV, phong.normalMatrix, phong.lightPosition,
                 phong.shadowIntensity, color=phong.color, clear_color=phong.clearColor)
    t_phong = time.perf_counter() - t0
    per_frame = (t_depth + t_screen) * (len(mvps) / samples) + t_phong
    npx = g_pos.shape[1] * g_pos.shape[2]
    return {"value": npx / per_frame / 1e6, "unit": "Mpix/s", "cores": cores, "threads": oracle.threads(),
            "nproc": os.cpu_count(), "kind": "port",
            "sample": f"{samples} of {len(mvps)} light samples (depth maps + screen pass) on the full "
                      f"{g_pos.shape[2]}x{g_pos.shape[1]} view, scaled x{len(mvps) // samples}, plus one "
                      f"Phong pass",
            "seconds": round(t_depth + t_screen + t_phong, 3)}


def diff_timing(gi, torch, steps, warmup):
    """Forward and forward+backward time of the differentiable soft-shadow
    pass on its own workload (16 samples, 1024^2 maps, 512^2 view, 50k
    triangles): HIP events around the library calls on the current stream;
    the backward recomputes the depth pass.  Matrices are resident."""
    import numpy as np

    from globalillumination_amd import passes
    from globalillumination_amd.diff import AreaLight, DiffScene

    wl = gi.scenes.config5()
    light, L, r = AreaLight.from_workload(wl)
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, passes.view_matrices(wl.camera)[2], wl.camera.width, wl.camera.height)
    scene = DiffScene(dm.indices, g, light)
    out = torch.empty((g.height, g.width), dtype=torch.float32, device=dm.positions.device)
    go = torch.randn((g.height, g.width), dtype=torch.float32, device=dm.positions.device,
                     generator=torch.Generator(device=dm.positions.device).manual_seed(5))
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        scene.forward(L, r, dm.positions, out=out)
        scene.backward_device(L, r, dm.positions, go)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for e in evs:
        for x in e:
            x.record(stream)
    torch.cuda.synchronize()
    for e in evs:
        e[0].record(stream)
        scene.forward(L, r, dm.positions, out=out)
        e[1].record(stream)
        scene.forward(L, r, dm.positions, out=out)
        scene.backward_device(L, r, dm.positions, go)
        e[2].record(stream)
    torch.cuda.synchronize()
    fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    fb = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    # Algorithmic bytes (lower bounds: the texels the taps read are not
    # counted).  Forward, per texel of the S maps: the 64-bit visibility key
    # written and read back, the FP64 depth written (24 B); per (map,
    # triangle): the vertices read twice (72 B), the 80-byte plane record
    # written and read back and its size flag (161 B); per pixel: position in,
    # output out (16 B).  Backward repeats the key and plane work and adds the
    # FP32 texel adjoint cleared, scatter-added and read (12 B per texel), the
    # key read by the texel adjoint (8 B) and the pixel adjoint's output
    # gradient read (4 B per pixel).
    S, msz, nt, npx = len(light.offsets), light.map_size ** 2, wl.mesh.ntris, g.width * g.height
    fwd_bytes = S * msz * 24 + S * nt * (72 + 161) + npx * 16
    bwd_bytes = S * msz * (16 + 12 + 8) + S * nt * (72 + 161) + npx * 20
    return {"workload": "differentiable soft shadows: 16 samples, 1024^2 maps, 512x512 view, "
                        f"{wl.mesh.ntris:,}-triangle displaced icosphere + ground",
            "fwd_ms": round(fwd, 4), "fwd_bwd_ms": round(fb, 4), "fwd_bwd_over_fwd": round(fb / fwd, 3),
            "steps": steps,
            "roofline": {"fwd_bytes": int(fwd_bytes), "fwd_GB/s": round(fwd_bytes / (fwd * 1e-3) / 1e9, 1),
                         "fwd_frac": round(fwd_bytes / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "fwd_bwd_bytes": int(fwd_bytes + bwd_bytes),
                         "fwd_bwd_GB/s": round((fwd_bytes + bwd_bytes) / (fb * 1e-3) / 1e9, 1),
                         "fwd_bwd_frac": round((fwd_bytes + bwd_bytes) / (fb * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def batch_timing(gi, torch, steps):
    """The 32-instance batch (maps 512^2-4096^2, views 640x480-1920x1080,
    5k-300k-triangle soups, 1-16 light samples, seed 4): the per-instance
    passes one after another, the batch's persistent launches, and the same
    launches replayed from a hipGraph, after the autotuner has picked the
    launch shape.  HIP events on the current stream; inputs resident."""
    import numpy as np

    from globalillumination_amd import passes
    from globalillumination_amd.batch import Batch, BatchInstance

    wls = gi.scenes.config4()
    soft = gi.SoftParams(pcf=True, kernelOrder=3, penumbra=1.0, inclusive=False)
    insts = []
    for wl in wls:
        dm = gi.DeviceMesh.from_mesh(wl.mesh)
        mvps = np.stack([passes.view_matrices(v)[2] for v in wl.lights])
        g = gi.render_gbuffer(dm, passes.view_matrices(wl.camera)[2], wl.camera.width, wl.camera.height)
        insts.append(BatchInstance(dm, mvps, wl.lights[0].width, wl.lights[0].width, g, soft))
    b = Batch(insts)
    tuning, tuned_ms = b.autotune(reps=2)
    graph = b.graph()
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def unbatched():
        for inst in insts:
            inst.run_unbatched()

    t_unb = timed(unbatched)
    t_run = timed(b.run)
    t_graph = timed(graph.launch)
    graph.close()
    b.close()
    npx = sum(wl.camera.width * wl.camera.height for wl in wls)
    # Algorithmic bytes of one batch (a lower bound: the texels the soft pass
    # gathers are not counted): every map written once (4 B per texel), the
    # triangle vertices read per map (36 B), position in and shadow out per
    # pixel (16 B)
    nbytes = sum(len(wl.lights) * (wl.lights[0].width ** 2 * 4 + wl.mesh.ntris * 36) +
                 wl.camera.width * wl.camera.height * 16 for wl in wls)
    return {"workload": "batched frames: 32 instances, maps 512^2-4096^2, views 640x480-1920x1080, "
                        "5k-300k-triangle soups, 1-16 light samples (seed 4)",
            "instances": len(wls), "view_pixels": int(npx), "light_samples": int(sum(len(w.lights) for w in wls)),
            "triangles": int(sum(w.mesh.ntris for w in wls)),
            "unbatched_ms": round(t_unb, 3), "batched_ms": round(t_run, 3), "graph_ms": round(t_graph, 3),
            "speedup_graph_vs_unbatched": round(t_unb / t_graph, 3),
            "graph_Mpix_per_s": round(npx / (t_graph * 1e-3) / 1e6, 1),
            "tuning": tuning, "autotune_ms": round(tuned_ms, 3), "steps": steps,
            "roofline": {"bytes": int(nbytes), "GB/s": round(nbytes / (t_graph * 1e-3) / 1e9, 1),
                         "frac": round(nbytes / (t_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def passes_timing(gi, torch, steps):
    """The other screen passes on one far-light frame: the area-light scene
    (100,822 triangles) lit from 40x further away with a 5-degree light
    frustum (window depths near 0.997, so PCSS and RBSSM find penumbrae),
    1920x1080 view, 2048^2 maps.  Hard-shadow techniques, PCSS (radius 64,
    kernel 15), RBSSM (radius 64, kernel 7, maxSearch 16), and the
    revectorization-based adaptive passes over the 9 samples of a 3 x 3
    light: one discontinuity layer, one level-0 node test, the final pass
    (binary and RPCF).  HIP events on the current stream; inputs resident."""
    import numpy as np

    from globalillumination_amd import passes

    wl = gi.scenes.config3(nsamples=9, sm=2048)
    wl.lights = [gi.scenes.View((120.0, 280.0, 80.0), (0.0, 0.0, 0.0), 2048, 2048, fov=5.0)] + \
        [gi.scenes.View(tuple(np.float32(v.eye) * np.float32(40.0)), v.at, 2048, 2048, fov=5.0) for v in wl.lights]
    W, H = wl.camera.width, wl.camera.height
    dm = gi.DeviceMesh.from_mesh(wl.mesh)
    g = gi.render_gbuffer(dm, passes.view_matrices(wl.camera)[2], W, H)
    mvps = np.stack([passes.view_matrices(v)[2] for v in wl.lights])
    maps = passes.render_shadow_map(dm, mvps, 2048, 2048, polygon_offset=(4.0, 20.0))
    MV, nm, lpos = passes.phong_uniforms(wl.camera, wl.lights[0].eye)
    out = torch.empty((H, W), dtype=torch.float32, device=maps.device)
    stream = torch.cuda.current_stream()

    def timed(fn):
        return round(event_ms(torch, fn, steps, stream), 4)

    res = {"workload": "far-light frame: 100,822-triangle displaced icosphere + ground, 1920x1080 view, "
                       "2048^2 maps (light 40x further, 5-degree frustum)", "steps": steps}
    # Algorithmic bytes per launch (`roofline` below): per foreground pixel the
    # G-buffer planes read and the output written, 8 B per background pixel,
    # and 4 B per unique texel (hard_footprint) of the taps every pixel
    # reads.  RBSM walks, PCSS/RBSSM's PCF taps (their extent depends on the
    # blocker search's result) and the revectorized passes' texels are not
    # counted: those figures are lower bounds.
    nfg = int((g.position[0] != 0).sum())
    roof = {}

    def note(key, ms, nbytes):
        gbs = nbytes / (ms * 1e-3) / 1e9
        roof[key] = {"bytes": int(nbytes), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    b0 = passes.biased(mvps[0])
    d0 = maps[0]
    tapsets = {"naive": ([0.0], False), "pcf": (tap_offsets(3, 1.0, False), False), "rbsm": ([0.0], True),
               "rpcf": (tap_offsets(3, 1.0, True), True), "rbsm_filtered": (tap_offsets(3, 1.0, True), True),
               "rbsm_conservative": ([0.0], True), "rpcf_conservative": (tap_offsets(3, 1.0, True), True)}
    for tech in ("naive", "pcf", "rbsm", "rpcf", "rbsm_filtered", "rbsm_conservative", "rpcf_conservative"):
        sp = gi.ShadowParams(lightMVP=mvps[0], shadowMapWidth=2048, shadowMapHeight=2048, technique=tech,
                             preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos,
                             depthThreshold=2.5e-5, maxSearch=16)
        res[f"hard_{tech}_ms"] = timed(lambda: passes.compute_hard_shadows(g, d0, sp, out=out))
        taps, nb = tapsets[tech]
        note(f"hard_{tech}", res[f"hard_{tech}_ms"],
             nfg * 28 + (W * H - nfg) * 8 + 4 * hard_footprint(torch, g, b0, 2048, 2048, taps, nb))
    # the 7 x 7 blocker search of PCSS and RBSSM: offsets w * (radius / 1024) / 3
    # in texture coordinates (2048-texel map), i.e. w * 2048 * (64/1024) / 3 texels
    search = [w * 2048.0 * (64.0 / 1024.0) / 3.0 for w in range(-3, 4)]
    search_texels = hard_footprint(torch, g, b0, 2048, 2048, search, False)
    pc = gi.PCSSParams(lightMVP=mvps[0], shadowMapWidth=2048, shadowMapHeight=2048, lightSourceRadius=64,
                       preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos)
    res["pcss_ms"] = timed(lambda: passes.render_soft_shadows(g, d0, pc, out=out))
    note("pcss", res["pcss_ms"], nfg * 28 + (W * H - nfg) * 8 + 4 * search_texels)
    rb = gi.RBSSMParams(lightMVP=mvps[0], shadowMapWidth=2048, shadowMapHeight=2048, lightSourceRadius=64,
                        preEvaluation=True, MV=MV, normalMatrix=nm, lightPosition=lpos, maxSearch=16,
                        depthThreshold=2.5e-5)
    res["rbssm_ms"] = timed(lambda: passes.render_rbssm(g, d0, rb, out=out))
    note("rbssm", res["rbssm_ms"], nfg * 28 + (W * H - nfg) * 8 + 4 * search_texels)
    samp = maps[1:].contiguous()
    biased = np.stack([passes.biased(m) for m in mvps[1:]])
    disc = passes.rev_discontinuity_layers(g)
    eyes = [tuple(float(x) for x in v.eye) for v in wl.lights[1:]]
    res["rev_discontinuity_ms"] = timed(
        lambda: passes.rev_discontinuity(g, samp, 0, biased[0], MV, nm, eyes[0], 0.25, disc))
    # position + normal in, one discontinuity byte out per pixel; the tap and
    # its four neighbours
    note("rev_discontinuity", res["rev_discontinuity_ms"],
         nfg * 25 + (W * H - nfg) * 5 + 4 * hard_footprint(torch, g, biased[0], 2048, 2048, [0.0], True))
    for i in range(1, 9):
        passes.rev_discontinuity(g, samp, i, biased[i], MV, nm, eyes[i], 0.25, disc)
    vis = torch.zeros((H, W), dtype=torch.float32, device=maps.device)
    img = torch.empty((H, W), dtype=torch.float32, device=maps.device)
    corners = [0, 2, 6, 8]
    _, count = passes.rev_quadtree_node_test(g, samp, corners, biased[corners], disc, 0, vis, image=img)
    count_dev = torch.empty((1,), dtype=torch.int32, device=maps.device)
    import ctypes

    from globalillumination_amd import _lib
    idx = (ctypes.c_int * 4)(*corners)
    gc = g.to_c(with_normals=False)
    m64 = passes._f(biased[corners], 64)

    def node_test():
        _lib.call("gi_rev_quadtree_node_test", ctypes.byref(gc), passes._dptr(samp), 9, 2048, 2048, idx, m64,
                  passes._dptr(disc), int(disc.shape[0]), 0, 1, passes._dptr(img), passes._dptr(vis),
                  passes._dptr(count_dev), passes._stream_handle(None))
    res["rev_node_test_ms"] = timed(node_test)
    # position in, four corner taps, four discontinuity bytes, image and
    # classification out (texels not counted)
    note("rev_node_test", res["rev_node_test_ms"], nfg * (12 + 4 + 4 + 4) + (W * H - nfg) * 12)
    common, trans = passes.adaptive_trans(mvps[1:], [1, 2, 1, 2, 1, 2, 1, 2, 1])
    pen = float(((vis == 0.5) & (g.position[0] != 0)).float().mean())
    res["rev_penumbra_pixel_fraction"] = round(pen, 4)
    for rp in (False, True):
        prm = gi.RevSoftParams(MV, nm, lpos, RPCF=rp, maxSearch=16, depthThreshold=2.5e-5)
        key = f"rev_soft{'_rpcf' if rp else ''}"
        res[f"{key}_ms"] = timed(
            lambda: passes.render_rev_adaptive(g, samp, common, trans, disc, vis, prm, out=out))
        # position + normal + classification in, output out (texels not counted)
        note(key, res[f"{key}_ms"], nfg * (24 + 4 + 4) + (W * H - nfg) * 8)
    res["roofline"] = roof
    return res


def main():
    args = _parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the shadow pipeline is measured per GPU "
              "(sharding by light sample is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.calibrate_copy:
        n = 256 << 20
        x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}))
        return

    import globalillumination_amd as gi
    from globalillumination_amd import passes

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload != "cfg3":
        raise SystemExit("only cfg3 (the metric's workload) is benchmarked here")
    wl = gi.scenes.config3()
    smw = wl.lights[0].width
    W, H = wl.camera.width, wl.camera.height
    dm = gi.DeviceMesh.from_mesh(wl.mesh, dev)
    mvps = np.stack([passes.view_matrices(v)[2] for v in wl.lights])
    S = mvps.shape[0]
    g = gi.render_gbuffer(dm, passes.view_matrices(wl.camera)[2], W, H)
    maps = torch.empty((S, smw, smw), dtype=torch.float32, device=dev)
    out = torch.empty((H, W), dtype=torch.float32, device=dev)
    params = gi.SoftParams(pcf=True, kernelOrder=3, penumbra=1.0, inclusive=False)
    common, trans = gi.soft_uniforms(mvps, wl.zbias)
    trans_dev = torch.from_numpy(trans).to(dev)
    mvps_dev = torch.from_numpy(np.ascontiguousarray(mvps, dtype=np.float32)).to(dev)
    # deferredShading: Phong with the centre light rotated 180 degrees about y
    phong = gi.PhongParams.from_views(wl.camera, wl.extra["center"], 0.25)
    rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def frame(ev=None):
        if ev:
            ev[0].record(stream)
        passes.render_shadow_map(dm, mvps_dev, smw, smw, out=maps, raster_events=(ev[3], ev[4]) if ev else None)
        if ev:
            ev[1].record(stream)
        passes.render_monte_carlo(g, maps, mvps, params, out=out, trans_dev=trans_dev)
        if ev:
            ev[2].record(stream)
        passes.shade_scene(g, out, phong, out=rgba)
        if ev:
            ev[5].record(stream)

    for _ in range(args.warmup):
        frame()
    torch.cuda.synchronize()
    # events: frame start, depth done, screen done, raster kernel begin / end, shading done
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(args.steps)]
    for e in evs:
        for x in e:
            x.record(stream)  # creates the underlying hipEvent
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        frame(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    depth_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    screen_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    raster_ms = float(np.mean([e[3].elapsed_time(e[4]) for e in evs]))
    phong_ms = float(np.mean([e[2].elapsed_time(e[5]) for e in evs]))
    value = W * H / (ms_step * 1e-3) / 1e6

    # Algorithmic bytes per launch.  Screen pass (k_soft): per foreground
    # pixel the G-buffer position (12 B) and the R32F output (4 B), per
    # background pixel the position x that identifies it (4 B) and the 0
    # written (4 B), plus every unique depth texel the taps gather (4 B
    # each).  Tile raster (k_raster): the maps written once (4 B/texel) + the
    # triangle vertices read per map (3 x 12 B per triangle).
    taps = tap_offsets(3, 1.0, False)
    # --no-footprint (profiling runs) leaves the texels uncounted: the
    # screen pass's algorithmic bytes, and every figure built on them, are null
    texels = None if args.no_footprint else touched_texels(gi, torch, g, mvps, smw, smw, taps)
    npx = W * H
    nfg = int((g.position[0] != 0).sum())
    screen_bytes = None if texels is None else nfg * (12 + 4) + (npx - nfg) * 8 + texels * 4
    raster_bytes = S * smw * smw * 4 + S * wl.mesh.ntris * 36
    # Phong: position, normal, facing and shadow in (28 B), RGBA out (16 B) per
    # foreground pixel; position x in and the clear colour out (20 B) per
    # background pixel
    phong_bytes = nfg * 48 + (npx - nfg) * 20
    screen_gbs = None if screen_bytes is None else screen_bytes / (screen_ms * 1e-3) / 1e9
    raster_gbs = raster_bytes / (raster_ms * 1e-3) / 1e9
    if raster_ms >= screen_ms:
        dom = {"kernel": "k_raster (tile raster of the depth pass)", "achieved": raster_gbs, "bytes": raster_bytes,
               "ms": raster_ms}
    else:
        dom = {"kernel": "k_soft (soft-shadow screen pass)", "achieved": screen_gbs, "bytes": screen_bytes,
               "ms": screen_ms}
    dom_name = "k_raster" if dom["kernel"].startswith("k_raster") else "k_soft"
    def rnd(x, n):
        return None if x is None else round(x, n)

    roof = {"bound": "hbm", "achieved": rnd(dom["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": rnd(None if dom["achieved"] is None else dom["achieved"] / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom_name),
            "traffic_source": str(TRAFFIC_SUMMARY.relative_to(ROOT)), "kernel": dom["kernel"],
            "algorithmic_bytes": None if dom["bytes"] is None else int(dom["bytes"]), "ms": round(dom["ms"], 4),
            "screen_pass": {"kernel": "k_soft", "ms": round(screen_ms, 4), "GB/s": rnd(screen_gbs, 1),
                            "taps_per_s": round(npx * S * len(taps) ** 2 / (screen_ms * 1e-3), -6),
                            "frac": rnd(None if screen_gbs is None else screen_gbs / HBM_PEAK_GBS, 4),
                            "bytes": screen_bytes, "unique_texels": texels, "traffic": measured_traffic("k_soft"),
                            "traffic_GB/s": (round(measured_traffic("k_soft") / (screen_ms * 1e-3) / 1e9, 1)
                                             if measured_traffic("k_soft") else None)},
            "depth_pass": {"ms": round(depth_ms, 4), "raster_kernel_ms": round(raster_ms, 4),
                           "raster_GB/s": round(raster_gbs, 1), "raster_bytes": int(raster_bytes)},
            "phong_pass": {"kernel": "k_phong", "ms": round(phong_ms, 4), "bytes": int(phong_bytes),
                           "GB/s": round(phong_bytes / (phong_ms * 1e-3) / 1e9, 1),
                           "frac": round(phong_bytes / (phong_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    # The reference's own soft-shadow PCF loop (AccurateSoftShadow.frag:101-119,
    # `<=`) takes 4 x 4 taps per sample where the workload names 3 x 3: the
    # screen pass in that mode, over the frame's maps, and the frame it gives.
    params16 = gi.SoftParams(pcf=True, kernelOrder=3, penumbra=1.0, inclusive=True)
    ev16 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    passes.render_monte_carlo(g, maps, mvps, params16, out=out, trans_dev=trans_dev)
    for a0, a1 in ev16:
        a0.record(stream)
        passes.render_monte_carlo(g, maps, mvps, params16, out=out, trans_dev=trans_dev)
        a1.record(stream)
    torch.cuda.synchronize()
    screen16_ms = float(np.mean([a0.elapsed_time(a1) for a0, a1 in ev16]))
    ref_loop = {"taps_per_sample": 16, "screen_ms": round(screen16_ms, 4),
                "taps_per_s": round(npx * S * 16 / (screen16_ms * 1e-3), -6),
                "frame_ms_events": round(depth_ms + screen16_ms + phong_ms, 4),
                "frame_Mpix_per_s_events": round(npx / ((depth_ms + screen16_ms + phong_ms) * 1e-3) / 1e6, 1)}
    diff = None if args.no_diff else diff_timing(gi, torch, args.diff_steps, 2)
    other = None if args.no_passes else passes_timing(gi, torch, args.passes_steps)
    rbsm = None if args.no_rbsm else config2_timing(gi, torch, args.passes_steps)
    batched = None if args.no_batch else batch_timing(gi, torch, args.batch_steps)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, mvps, g.position.cpu().numpy(), g.normal.cpu().numpy(), g.facing.cpu().numpy(),
                           phong, smw, args.cpu_samples, gi)
    line = {"metric": "shaded Mpix/s (1080p, 64-sample soft shadow)", "value": round(value, 2),
            "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "area-light soft shadows: 64 samples, 2048^2 FP32 depth maps, 1920x1080 "
                                   "view, 3x3 PCF per sample, deferred Phong, 100,822-triangle displaced "
                                   "icosphere + ground",
                       "global_batch": 1, "samples": S, "shadow_map": smw, "view": [W, H],
                       "triangles": wl.mesh.ntris},
            "roofline": roof, "cpu_baseline": cpu, "reference_pcf_loop": ref_loop, "differentiable": diff,
            "batched": batched, "other_passes": other, "rbsm_workload": rbsm}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
