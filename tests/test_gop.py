"""Picture-type decisions of the encoder driver (ironav_amd/csrc/gop.c) on the CPU.

Pinned by the reference's own encoder tests, restated with their inputs: a scene cut at the sixth of
ten 96x64 pictures starts an IDR picture there (IDR pictures at 0 and 5,
crates/ironav-codec/src/h264/enc/mod.rs:1071-1107); IDR pictures every 5 pictures with one forced at
picture 3 give IDR pictures at 0, 3 and 8 without B pictures (enc/mod.rs:1110-1136). The scene-cut
measure itself against an independent numpy statement of change.rs:57-89 (the same sums for any
thread count), the low-resolution luma against numpy block means.
"""
import numpy as np
import pytest

from ironav_amd import abi, h264w


def ref_scene(w, h, t):
    """The luma of the reference tests' `scene(w, h, t)` (enc/mod.rs:745-763): a gradient, a bright
    band moving left by 3 samples per picture, a dark one moving right by 2, and LCG noise in -4..4."""
    seed = (7 + t * 1000) & 0xFFFFFFFF
    y = np.zeros((h, w), dtype=np.uint8)
    for r in range(h):
        for c in range(w):
            seed = (seed * 1103515245 + 12345) & 0xFFFFFFFF
            noise = ((seed >> 16) % 9) - 4
            v = c * 160 // w + r * 60 // h + 20
            if (c + 3 * t) % w < w // 4 and h // 5 < r < h // 2:
                v = 220 - ((c + r) % 7) * 10
            u = (c + w - (2 * t) % w) % w
            if w // 2 < u < w // 2 + w // 6 and r > h // 2 + 4:
                v = 30 + ((r * 5) % 40)
            y[r, c] = min(max(v + noise, 0), 255)
    return y


def test_scene_cut_starts_an_idr_picture():
    w, h = 96, 64
    kinds = []
    with h264w.Gop(w, h) as g:
        for t in range(10):
            y = ref_scene(w, h, t)
            if t >= 5:
                y = (255 - (np.arange(w * h) // w) * 3).astype(np.uint8).reshape(h, w)
            kinds.append(g.next(y))
    assert [t for t, k in enumerate(kinds) if k == abi.PIC_IDR] == [0, 5]


def test_keyint_and_forced_idr():
    w, h = 96, 64
    kinds = []
    with h264w.Gop(w, h, keyint=5) as g:
        for t in range(12):
            kinds.append(g.next(ref_scene(w, h, t), force_idr=(t == 3)))
    assert [t for t, k in enumerate(kinds) if k == abi.PIC_IDR] == [0, 3, 8]


def np_change(prev, cur):
    lh, lw = cur.shape
    inter = intra = 0
    for by in range(0, lh - 7, 8):
        for bx in range(0, lw - 7, 8):
            blk = cur[by:by + 8, bx:bx + 8].astype(np.int64)
            mean = int(blk.sum()) // 64
            intra += int(np.abs(blk - mean).sum())
            best = None
            for dy in range(-4, 5):
                for dx in range(-4, 5):
                    oy, ox = by + dy, bx + dx
                    if ox < 0 or oy < 0 or ox + 8 > lw or oy + 8 > lh:
                        continue
                    s = int(np.abs(blk - prev[oy:oy + 8, ox:ox + 8]).sum())
                    best = s if best is None else min(best, s)
            inter += best
    return inter, intra


@pytest.mark.parametrize("lw,lh", [(24, 16), (37, 29), (64, 40), (8, 8), (7, 20)])
def test_change_matches_numpy(lw, lh):
    rng = np.random.default_rng(lw * 100 + lh)
    prev = rng.integers(0, 256, (lh, lw)).astype(np.int32)
    cur = np.roll(prev, (2, -3), axis=(0, 1)) + rng.integers(-6, 7, (lh, lw)).astype(np.int32)
    want = np_change(prev, cur)
    for threads in (1, 3, 8):
        assert h264w.change(prev, cur, threads) == want, threads


def test_lowres_is_the_block_mean():
    rng = np.random.default_rng(3)
    y = rng.integers(0, 256, (48, 80)).astype(np.uint8)
    want = y.reshape(12, 4, 20, 4).astype(np.int32).sum(axis=(1, 3)) >> 4
    assert np.array_equal(h264w.lowres(y), want)
