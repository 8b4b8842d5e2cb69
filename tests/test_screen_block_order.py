"""The screen passes' workgroup-to-block order (pixel_of_block<k> and
soft_block_pixel in globalillumination_amd/csrc/screen_kernels.hpp),
restated: for every view size and every strip count k per XCD, the 8-XCD
round-robin of workgroups followed by the strip-major order must visit each
16 x 16 block exactly once.  With one strip per XCD each XCD's blocks lie in
at most two neighbouring strips; with several, each XCD's blocks spread
across the view.  The soft pass's 64 x 4 blocks of 16 x 4 waves must cover
every pixel exactly once."""
import numpy as np
import pytest

BLOCK = 16


def block_of(b, bx_n, by_n, spx=1):
    nblocks = bx_n * by_n
    per, rem = divmod(nblocks, 8)
    xcd, slot = b % 8, b // 8
    lin = xcd * (per + 1) + slot if xcd < rem else rem * (per + 1) + (xcd - rem) * per + slot
    if lin >= nblocks:
        lin = b
    nstrips = 8 * spx
    sw = (bx_n + nstrips - 1) // nstrips
    acc = 0
    for o in range(nstrips):
        s = (o % spx) * 8 + o // spx  # XCD k drains strips k, k + 8, ...
        x0 = s * sw
        wk = 0 if x0 >= bx_n else min(sw, bx_n - x0)
        if lin < acc + wk * by_n:
            q = lin - acc
            return x0 + q % wk, q // wk, xcd, s
        acc += wk * by_n
    raise AssertionError("position past the last strip")


SIZES = [(1, 1), (16, 16), (17, 5), (33, 17), (100, 257), (320, 200), (480, 270), (512, 512), (640, 480),
         (800, 600), (1280, 720), (1920, 1080), (4000, 3), (3, 4000)]


@pytest.mark.parametrize("spx", [1, 2, 4, 8])
@pytest.mark.parametrize("w,h", SIZES)
def test_block_order_is_a_permutation(w, h, spx):
    bx_n, by_n = (w + BLOCK - 1) // BLOCK, (h + BLOCK - 1) // BLOCK
    seen = set()
    strips = {}
    for b in range(bx_n * by_n):
        x, y, xcd, k = block_of(b, bx_n, by_n, spx)
        assert 0 <= x < bx_n and 0 <= y < by_n
        seen.add((x, y))
        strips.setdefault(xcd, set()).add(k)
    assert len(seen) == bx_n * by_n
    if spx == 1:
        for ks in strips.values():
            assert max(ks) - min(ks) <= 1


def test_full_hd_spreads_each_xcd_with_four_strips():
    # 120 x 68 blocks, 32 strips of 4 columns (30 non-empty): every XCD holds
    # blocks of at least 3 strips, spanning more than half the view's width
    xs, ks = {}, {}
    for b in range(120 * 68):
        x, y, xcd, k = block_of(b, 120, 68, 4)
        xs.setdefault(xcd, []).append(x)
        ks.setdefault(xcd, set()).add(k)
    for xcd in range(8):
        assert len(ks[xcd]) >= 3
        assert max(xs[xcd]) - min(xs[xcd]) > 60


def test_full_hd_gives_each_xcd_one_whole_strip():
    # 120 x 68 blocks: 8 strips of 15 columns, 1,020 blocks each
    for b in range(120 * 68):
        x, y, xcd, k = block_of(b, 120, 68)
        assert k == xcd and 15 * xcd <= x < 15 * xcd + 15


def soft_block_pixels(w, h, bw=64, wx=16, spx=4):
    """Pixels of every thread of the soft pass's grid (soft_block_pixel)."""
    bh, wy = 256 // bw, 64 // wx
    bx_n, by_n = (w + bw - 1) // bw, (h + bh - 1) // bh
    t = np.arange(256)
    lane, wave = t & 63, t >> 6
    ox = (wave % (bw // wx)) * wx + lane % wx
    oy = (wave // (bw // wx)) * wy + lane // wx
    xs, ys = [], []
    for b in range(bx_n * by_n):
        bx, by, _, _ = block_of(b, bx_n, by_n, spx)
        xs.append(bx * bw + ox)
        ys.append(by * bh + oy)
    return np.concatenate(xs), np.concatenate(ys)


@pytest.mark.parametrize("w,h", [(1, 1), (17, 5), (100, 257), (640, 480), (1920, 1080), (4000, 3), (3, 4000)])
def test_soft_blocks_cover_each_pixel_once(w, h):
    x, y = soft_block_pixels(w, h)
    inside = (x < w) & (y < h)
    k = y[inside].astype(np.int64) * w + x[inside]
    assert len(k) == w * h
    assert len(np.unique(k)) == w * h
