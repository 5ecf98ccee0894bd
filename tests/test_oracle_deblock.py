"""The oracle's loop filter (h264/deblock.rs:608-695 as the encoder runs it, enc/mod.rs:653-654).

The reference holds no golden vectors for the filter: its only test compares its SIMD and scalar
edges with each other (h264/deblock.rs:929-963), and the decoder fixtures that would pin it need a
decoder. So the oracle is checked against what H.264 8.7 fixes, chosen so that a wrong stride, a
swapped p/q side, a wrong edge position or a wrong threshold fails one of them:
  - identities: constant planes; every strength 0 (inter, one vector, one reference, no
    coefficients); qp below 16 (alpha = 0); steps of alpha or more;
  - symmetries: exchanging the two sides of an edge mirrors the result; transposing a picture (and
    its records) transposes the result;
  - the strong filter's outputs are convex combinations of the inputs: they stay within the step;
  - reach: chroma changes only the sample either side of the chroma edges (x mod 4 in {0, 3}), luma
    at most three samples either side of the 4x4 grid (and of the 8x8 grid in 8x8-transform MBs).
The GPU kernel is then checked byte for byte against this oracle (tests/test_gpu_deblock.py).
"""
import numpy as np
import pytest

from ironav_amd import abi

I4 = 1  # IA_KIND_I4X4; 5 is IA_KIND_INTER


def records(mb_w, mb_h, kind=5, qp=28, nnz=0, mv=(0, 0), ref=0, t8x8=0):
    m = np.zeros(mb_w * mb_h, dtype=abi.MBINFO_DTYPE)
    m["slice"] = 1
    m["kind"] = kind
    m["qp"] = qp
    m["t8x8"] = t8x8
    m["nnz"] = nnz
    m["mv"][:, :, 0] = mv[0]
    m["mv"][:, :, 1] = mv[1]
    m["ref_idx"] = ref if kind == 5 else -1
    m["ref_id"] = ref + 1 if kind == 5 else -1
    return m


def planes(mb_w, mb_h, rng, lo=0, hi=256):
    return (rng.integers(lo, hi, (mb_h * 16, mb_w * 16), dtype=np.uint8),
            rng.integers(lo, hi, (mb_h * 8, mb_w * 8), dtype=np.uint8),
            rng.integers(lo, hi, (mb_h * 8, mb_w * 8), dtype=np.uint8))


def test_identities(oracle):
    rng = np.random.default_rng(1)
    w, h = 4, 3
    # constant planes under the strongest filtering (intra everywhere)
    y, cb, cr = (np.full((h * 16, w * 16), 97, np.uint8), np.full((h * 8, w * 8), 31, np.uint8),
                 np.full((h * 8, w * 8), 200, np.uint8))
    for got, want in zip(oracle.deblock_picture(y, cb, cr, records(w, h, kind=I4, qp=40), w, h), (y, cb, cr)):
        assert np.array_equal(got, want)
    # strength 0 everywhere: noise is left alone
    p = planes(w, h, rng, 100, 110)
    for got, want in zip(oracle.deblock_picture(*p, records(w, h, kind=5, qp=40, mv=(3, -7)), w, h), p):
        assert np.array_equal(got, want)
    # qp below 16: alpha is 0
    for got, want in zip(oracle.deblock_picture(*p, records(w, h, kind=I4, qp=15), w, h), p):
        assert np.array_equal(got, want)
    # 4x4 checkerboard of 100 / 130: every edge steps by 30, at least alpha at qp 30 (luma alpha 25,
    # chroma qp 29: alpha 22), so nothing changes; at qp 40 (alpha 80) the edges are filtered
    cy = (100 + (np.indices((h * 16, w * 16)) // 4).sum(0) % 2 * 30).astype(np.uint8)
    cc = (100 + (np.indices((h * 8, w * 8)) // 2).sum(0) % 2 * 30).astype(np.uint8)
    for got, want in zip(oracle.deblock_picture(cy, cc, cc, records(w, h, kind=I4, qp=30), w, h), (cy, cc, cc)):
        assert np.array_equal(got, want)
    got = oracle.deblock_picture(cy, cc, cc, records(w, h, kind=I4, qp=40), w, h)
    assert not np.array_equal(got[0], cy) and not np.array_equal(got[1], cc)


@pytest.mark.parametrize("kind,nnz", [(2, 0), (5, 3)])
def test_side_exchange_mirrors(oracle, kind, nnz):
    """Two flat macroblocks side by side (a | b) with 8x8 transforms: the only inner edges are 8
    samples inside each, out of reach of the macroblock edge's changes (3 samples), and flat, so only
    the macroblock edge changes anything. Exchanging a and b mirrors the result left-right (the
    filter's p and q formulas are mirror images, rounding included)."""
    w, h = 2, 1
    rec = records(w, h, kind=kind, qp=36, nnz=nnz, t8x8=1)
    for a, b in [(90, 100), (100, 90), (40, 47), (200, 180), (10, 13)]:
        def pic(u, v):
            y = np.empty((16, 32), np.uint8)
            y[:, :16], y[:, 16:] = u, v
            c = np.empty((8, 16), np.uint8)
            c[:, :8], c[:, 8:] = u, v
            return y, c.copy(), c.copy()
        ya, cba, _ = oracle.deblock_picture(*pic(a, b), rec, w, h)
        yb, cbb, _ = oracle.deblock_picture(*pic(b, a), rec, w, h)
        assert np.array_equal(ya, yb[:, ::-1]), (a, b)
        assert np.array_equal(cba, cbb[:, ::-1]), (a, b)
        # reach: luma within 3 samples of x = 16, chroma within 1 of x = 8; outputs stay within [a, b]
        changed = np.nonzero((ya != pic(a, b)[0]).any(0))[0]
        assert changed.size and changed.min() >= 13 and changed.max() <= 18
        assert ya.min() >= min(a, b) and ya.max() <= max(a, b)
        cch = np.nonzero((cba != pic(a, b)[1]).any(0))[0]
        assert set(cch.tolist()) <= {7, 8}


def ramp_pictures(rng):
    """Random sample lines across the macroblock edge of two macroblocks (every row the same). At the
    inner 8x8 edges (luma x 8 and 24, chroma x 4 and 12) the first sample pair inside steps by 50 or
    more, beyond every beta, so those edges stay unfiltered."""
    line = np.empty(32, np.int64)
    line[:8], line[24:] = 5, 250
    line[10:22] = rng.integers(90, 130) + rng.integers(-6, 7, 12)
    line[16:22] += rng.integers(-12, 13)
    line[8], line[9], line[22], line[23] = 220, 40, 40, 220
    line = np.clip(line, 0, 255).astype(np.uint8)
    y = np.repeat(line[None, :], 16, 0)
    c = np.repeat(np.concatenate([np.full(4, 5), line[8:24:2], np.full(4, 250)]).astype(np.uint8)[None, :], 8, 0)
    return y, c


@pytest.mark.parametrize("qp", [26, 36, 46])
def test_strong_edges_mirror(oracle, qp):
    """Intra 8x8-transform macroblocks (strength 4 at their edge, no inner edge within reach): the
    filtered picture of the mirrored lines is the mirror of the filtered picture. The strength-4
    updates of p and q are mirror images of each other, rounding included, so a wrong neighbour or
    weight on one side fails."""
    rng = np.random.default_rng(qp)
    rec = records(2, 1, kind=2, qp=qp, t8x8=1)
    for _ in range(300):
        y, c = ramp_pictures(rng)
        a = oracle.deblock_picture(y, c, c, rec, 2, 1)
        b = oracle.deblock_picture(y[:, ::-1].copy(), c[:, ::-1].copy(), c[:, ::-1].copy(), rec, 2, 1)
        for u, v in zip(a, b):
            assert np.array_equal(u, v[:, ::-1]), y[0]


@pytest.mark.parametrize("qp", [30, 40, 50])
def test_normal_edges_conserve(oracle, qp):
    """Inter 8x8-transform macroblocks with coefficients (strength 2): the filter moves p0 and q0 by
    +delta and -delta (no clipping at these levels), so p0 + q0 is conserved on every line, luma and
    chroma, and only p1..q1 change; with the inner edges unfiltered, nothing else changes."""
    rng = np.random.default_rng(qp)
    rec = records(2, 1, kind=5, qp=qp, nnz=2, t8x8=1)
    moved = 0
    for _ in range(300):
        y, c = ramp_pictures(rng)
        gy, gc, _ = oracle.deblock_picture(y, c, c, rec, 2, 1)
        assert np.array_equal(gy[:, 15].astype(int) + gy[:, 16], y[:, 15].astype(int) + y[:, 16])
        assert np.array_equal(gc[:, 7].astype(int) + gc[:, 8], c[:, 7].astype(int) + c[:, 8])
        changed = np.nonzero((gy != y).any(0))[0]
        assert set(changed.tolist()) <= {14, 15, 16, 17}
        assert set(np.nonzero((gc != c).any(0))[0].tolist()) <= {7, 8}
        moved += changed.size > 0
    assert moved > 50


def test_transpose(oracle):
    """Two flat macroblocks side by side (a | b) and their transpose, one above the other: each has
    one active macroblock edge, processed with the same inner edges after it; the filtered pictures
    are transposes of each other, in luma and chroma (the transpose exchanges the roles of the
    vertical and the horizontal edges, so a wrong stride or edge position fails)."""
    for kind, nnz, t8 in [(1, 0, 0), (5, 1, 0), (2, 0, 1)]:
        y = np.zeros((16, 32), np.uint8)
        y[:, :16], y[:, 16:] = 120, 131
        c = np.zeros((8, 16), np.uint8)
        c[:, :8], c[:, 8:] = 60, 66
        a = oracle.deblock_picture(y, c, c, records(2, 1, kind=kind, qp=38, nnz=nnz, t8x8=t8), 2, 1)
        b = oracle.deblock_picture(y.T.copy(), c.T.copy(), c.T.copy(), records(1, 2, kind=kind, qp=38, nnz=nnz,
                                                                           t8x8=t8), 1, 2)
        for u, v in zip(a, b):
            assert np.array_equal(u, v.T)
        assert not np.array_equal(a[0], y) and not np.array_equal(a[1], c)


def test_reach_on_noise(oracle):
    """A picture whose columns are constant: the vertical edges change columns, and a horizontal edge
    changes only columns where they left the two sides different. So chroma changes only at x mod 4
    in {0, 3} (the sample either side of its vertical edges), and with 8x8 transforms everywhere luma
    columns 3 + 8k and 4 + 8k (out of reach of the 8x8 grid's edges) never change."""
    rng = np.random.default_rng(5)
    w, h = 5, 4
    y = np.repeat(rng.integers(60, 90, (1, w * 16), dtype=np.uint8), h * 16, 0)
    c = np.repeat(rng.integers(60, 90, (1, w * 8), dtype=np.uint8), h * 8, 0)
    for kind in (1, 5):
        for t8 in (0, 1):
            rec = records(w, h, kind=kind, qp=44, nnz=2, t8x8=t8)
            gy, gc, _ = oracle.deblock_picture(y, c, c, rec, w, h)
            cols = np.nonzero((gc != c).any(0))[0]
            assert cols.size and set((cols % 4).tolist()) <= {0, 3}
            ycols = np.nonzero((gy != y).any(0))[0]
            assert ycols.size
            if t8:
                assert not set((ycols % 8).tolist()) & {3, 4}
