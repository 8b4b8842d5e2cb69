"""gi_depth_pass_noalloc (the depth pass on the library's static device
workspace) against gi_depth_pass with a caller workspace, through the C ABI:
bit-identical maps on two calls in a row, the second smaller than the first
(the static array then holds the first call's lists), and GI_ENOSPACE for a
workload whose workspace exceeds the array."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _depth(gi, name, mesh, mvps, w, h):
    from globalillumination_amd import _lib

    lib = _lib.lib()
    dev = torch.device("cuda")
    pos = torch.from_numpy(np.ascontiguousarray(mesh.positions, np.float32)).to(dev)
    idx = torch.from_numpy(np.ascontiguousarray(mesh.indices, np.int32)).to(dev)
    m = torch.from_numpy(np.ascontiguousarray(mvps, np.float32).reshape(-1, 16)).to(dev)
    n = m.shape[0]
    out = torch.full((n, h, w), -1.0, dtype=torch.float32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (ctypes.c_void_p(pos.data_ptr()), pos.shape[0], ctypes.c_void_p(idx.data_ptr()), idx.shape[0],
            ctypes.c_void_p(m.data_ptr()), n, w, h, None, ctypes.c_void_p(out.data_ptr()))
    if name == "gi_depth_pass":
        nbytes = lib.gi_depth_pass_workspace_size(idx.shape[0], n, w, h)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        rc = lib.gi_depth_pass(*args, ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(nbytes), stream)
    else:
        rc = lib.gi_depth_pass_noalloc(*args, stream)
    torch.cuda.synchronize()
    assert rc == 0, rc
    return out.cpu().numpy()


def test_noalloc_depth_pass_matches_workspace_pass(gi, cuda):
    from globalillumination_amd import _lib, passes

    lib = _lib.lib()
    assert lib.gi_depth_pass_noalloc_max_bytes() == 64 << 20
    # first call: the differentiable workload's maps (16 x 1024^2, 50k-triangle icosphere + ground)
    wl = gi.scenes.config5()
    mvps = np.stack([passes.view_matrices(v)[2] for v in wl.lights])
    need = lib.gi_depth_pass_workspace_size(wl.mesh.ntris, len(mvps), 1024, 1024)
    print(f"call 1: {wl.mesh.ntris} triangles, {len(mvps)} maps of 1024^2, workspace {need / 2**20:.1f} MiB",
          flush=True)
    assert need <= 64 << 20
    a = _depth(gi, "gi_depth_pass", wl.mesh, mvps, 1024, 1024)
    b = _depth(gi, "gi_depth_pass_noalloc", wl.mesh, mvps, 1024, 1024)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a < 1.0).mean() > 0.2
    # second call, smaller: 3 maps of 500 x 300 of the config-1 sphere
    wl1 = gi.scenes.config1()
    m1 = np.stack([passes.view_matrices(v)[2] for v in wl1.lights] * 3)
    print(f"call 2: {wl1.mesh.ntris} triangles, 3 maps of 500x300", flush=True)
    a = _depth(gi, "gi_depth_pass", wl1.mesh, m1, 500, 300)
    b = _depth(gi, "gi_depth_pass_noalloc", wl1.mesh, m1, 500, 300)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a < 1.0).mean() > 0.05


def test_noalloc_depth_pass_rejects_oversize(gi, cuda):
    """64 maps of 2048^2 with 100k triangles need ~470 MiB: rejected before
    any launch (the pointers are never read)."""
    from globalillumination_amd import _lib

    lib = _lib.lib()
    assert lib.gi_depth_pass_workspace_size(100_822, 64, 2048, 2048) > 64 << 20
    fake = ctypes.c_void_p(256)
    rc = lib.gi_depth_pass_noalloc(fake, 3, fake, 100_822, fake, 64, 2048, 2048, None, fake, None)
    assert rc == _lib.GI_ENOSPACE
    assert lib.gi_depth_pass_noalloc(fake, 3, fake, 10, fake, 0, 64, 64, None, fake, None) == _lib.GI_EINVAL
