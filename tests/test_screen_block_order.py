"""The screen passes' workgroup-to-block order (pixel_of_block in
globalillumination_amd/csrc/screen_kernels.hpp), restated: for every view
size, the 8-XCD round-robin of workgroups followed by the strip-major order
must visit each 16 x 16 block exactly once, and each XCD's blocks must lie in
at most two neighbouring strips."""
import pytest

BLOCK = 16


def block_of(b, bx_n, by_n):
    nblocks = bx_n * by_n
    per, rem = divmod(nblocks, 8)
    xcd, slot = b % 8, b // 8
    lin = xcd * (per + 1) + slot if xcd < rem else rem * (per + 1) + (xcd - rem) * per + slot
    if lin >= nblocks:
        lin = b
    strip = (bx_n + 7) // 8
    nstrips = (bx_n + strip - 1) // strip
    k = min(lin // (strip * by_n), nstrips - 1)
    q = lin - k * strip * by_n
    wk = min(strip, bx_n - k * strip)
    return k * strip + q % wk, q // wk, xcd, k


@pytest.mark.parametrize("w,h", [(1, 1), (16, 16), (17, 5), (33, 17), (100, 257), (320, 200), (480, 270), (512, 512),
                                 (640, 480), (800, 600), (1280, 720), (1920, 1080), (4000, 3), (3, 4000)])
def test_block_order_is_a_permutation(w, h):
    bx_n, by_n = (w + BLOCK - 1) // BLOCK, (h + BLOCK - 1) // BLOCK
    seen = set()
    strips = {}
    for b in range(bx_n * by_n):
        x, y, xcd, k = block_of(b, bx_n, by_n)
        assert 0 <= x < bx_n and 0 <= y < by_n
        seen.add((x, y))
        strips.setdefault(xcd, set()).add(k)
    assert len(seen) == bx_n * by_n
    for ks in strips.values():
        assert max(ks) - min(ks) <= 1


def test_full_hd_gives_each_xcd_one_whole_strip():
    # 120 x 68 blocks: 8 strips of 15 columns, 1,020 blocks each
    for b in range(120 * 68):
        x, y, xcd, k = block_of(b, 120, 68)
        assert k == xcd and 15 * xcd <= x < 15 * xcd + 15
