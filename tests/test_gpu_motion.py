"""GPU parity of the half-sample planes and the exhaustive search, through the C ABI, against the oracle.

Bit-exact: planes byte for byte; full search the same vector and SAD for every partition of every
macroblock and reference. At the 1080p workload size the oracle's brute force would take minutes,
so there a seeded sample of macroblocks is checked against a numpy brute force, and every returned
SAD is recomputed at its returned vector.
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Analyser, Picture

pytestmark = pytest.mark.gpu


def frames(w, h, n, seed, **kw):
    spec = clipgen.ClipSpec(w, h, n, seed, **kw)
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


@pytest.mark.parametrize("w,h", [(16, 16), (200, 120), (1280, 720), (1920, 1080), (3840, 2160)])
def test_halfpel_planes_bit_exact(oracle, w, h):
    pics = frames(w, h, 1, 21)
    with Analyser(w, h, nrefs=1) as an:
        an.set_reference(0, pics[0])
        got = an.read_halfpel(0)
    want = oracle.halfpel_build(pics[0].y.astype(np.uint16), 8)
    assert got.shape == want.shape
    for k in range(4):
        assert np.array_equal(got[k], want[k].astype(np.uint8)), f"plane {'GBHJ'[k]}"


@pytest.mark.parametrize("w,h,rng,nref", [(128, 96, 16, 3), (208, 112, 32, 2), (64, 48, 5, 1), (320, 256, 64, 1),
                                          (16, 16, 32, 1), (96, 32, 1, 2)])
def test_fullsearch_matches_oracle(oracle, w, h, rng, nref):
    pics = frames(w, h, nref + 1, 100 + rng, max_motion=rng // 2 + 1)
    with Analyser(w, h, nrefs=nref, search_mode=abi.SEARCH_FULL, search_range=rng) as an:
        for r in range(nref):
            an.set_reference(r, pics[r])
        an.upload(pics[nref])
        an.motion_search(list(range(nref)))
        got = an.read_motion(nref)
    want = oracle.fullsearch(pics[nref].y, [p.y for p in pics[:nref]], an.mb_w, an.mb_h, rng)
    assert np.array_equal(got["sad"], want["sad"])
    assert np.array_equal(got["mv"], want["mv"])


@pytest.mark.parametrize("rng", [32, 64])
def test_fullsearch_variants_agree(oracle, rng):
    """Every tile shape against the oracle; at +-64 the window (129 columns) is wider than the
    128-thread tiles, so a lane visits two columns of one row."""
    pics = frames(176, 144, 2, 77, max_motion=20)
    results = []
    with Analyser(176, 144, nrefs=1, search_mode=abi.SEARCH_FULL, search_range=rng) as an:
        an.set_reference(0, pics[0])
        an.upload(pics[1])
        from ironav_amd.h264a import fullsearch_variants
        for v in range(len(fullsearch_variants())):
            an.set_fullsearch_variant(v)
            an.motion_search([0])
            results.append(an.read_motion(1))
    want = oracle.fullsearch(pics[1].y, [pics[0].y], an.mb_w, an.mb_h, rng)
    for v, got in enumerate(results):
        assert np.array_equal(got["sad"], want["sad"]), v
        assert np.array_equal(got["mv"], want["mv"]), v


def test_fullsearch_ties_resolve_in_raster_order(oracle):
    """A flat reference with one textured square: most displacements tie, and each partition's winner
    must be the first in raster order of (dy, dx), for every tile shape, at +-64."""
    w, h = 192, 160
    y = np.full((h, w), 90, dtype=np.uint8)
    y[64:80, 96:112] = np.arange(256, dtype=np.uint8).reshape(16, 16)
    c = np.full((h // 2, w // 2), 128, dtype=np.uint8)
    ref, src = Picture(y, c, c), Picture(np.roll(y, (3, -5), axis=(0, 1)), c, c)
    want = oracle.fullsearch(src.y, [ref.y], w // 16, h // 16, 64)
    with Analyser(w, h, nrefs=1, search_mode=abi.SEARCH_FULL, search_range=64) as an:
        an.set_reference(0, ref)
        an.upload(src)
        from ironav_amd.h264a import fullsearch_variants
        for v in range(len(fullsearch_variants())):
            an.set_fullsearch_variant(v)
            an.motion_search([0])
            got = an.read_motion(1)
            assert np.array_equal(got["sad"], want["sad"]), v
            assert np.array_equal(got["mv"], want["mv"]), v


def _brute_one(src, refpad, mbx, mby, pw, ph, rng):
    x0, y0 = mbx * 16, mby * 16
    dxs = np.arange(max(-rng, -(x0 + 24)), min(rng, pw - x0 + 8) + 1)
    dys = np.arange(max(-rng, -(y0 + 24)), min(rng, ph - y0 + 8) + 1)
    s = src[y0:y0 + 16, x0:x0 + 16].astype(np.int32)
    # all candidate windows at once: (ndy, ndx, 16, 16)
    win = np.lib.stride_tricks.sliding_window_view(refpad, (16, 16))
    cand = win[y0 + dys[0] + 48:y0 + dys[-1] + 49, x0 + dxs[0] + 48:x0 + dxs[-1] + 49].astype(np.int32)
    a = np.abs(cand - s)
    out = []
    for (x, y, w, h) in abi.PARTITIONS:
        sad = a[:, :, y * 4:(y + h) * 4, x * 4:(x + w) * 4].sum(axis=(2, 3))
        k = int(np.argmin(sad))  # first minimum in raster order of (dy, dx)
        iy, ix = divmod(k, sad.shape[1])
        out.append((int(sad[iy, ix]), int(dxs[ix]) * 4, int(dys[iy]) * 4))
    return out


def test_fullsearch_1080p_workload_sampled():
    """The 1080p, +-32, 3-reference workload: a seeded sample of 48 macroblocks (with the picture's
    corners and edges) against a numpy brute force, plus every SAD recomputed at its vector."""
    pics = frames(1920, 1080, 4, 1080, max_motion=12)
    with Analyser(1920, 1080, nrefs=3, search_mode=abi.SEARCH_FULL, search_range=32) as an:
        for r in range(3):
            an.set_reference(r, pics[2 - r])
        an.upload(pics[3])
        an.motion_search([0, 1, 2])
        got = an.read_motion(3)
        mb_w, mb_h, pw, ph = an.mb_w, an.mb_h, an.pw, an.ph
    rs = np.random.default_rng(1)
    sample = {(0, 0), (mb_w - 1, 0), (0, mb_h - 1), (mb_w - 1, mb_h - 1), (mb_w // 2, 0), (0, mb_h // 2)}
    while len(sample) < 48:
        sample.add((int(rs.integers(mb_w)), int(rs.integers(mb_h))))
    for r in range(3):
        refpad = np.pad(pics[2 - r].y.astype(np.int32), 48, mode="edge")
        for mbx, mby in sorted(sample):
            want = _brute_one(pics[3].y, refpad, mbx, mby, pw, ph, 32)
            g = got[r, mby * mb_w + mbx]
            for p in range(abi.NPART):
                assert (int(g["sad"][p]), int(g["mv"][p][0]), int(g["mv"][p][1])) == want[p], (r, mbx, mby, p)
    # every SAD equals the SAD at its own vector (all macroblocks, the 16x16 partition)
    src = pics[3].y.astype(np.int32)
    for r in range(3):
        refpad = np.pad(pics[2 - r].y.astype(np.int32), 48, mode="edge")
        mv = got[r]["mv"][:, 0] // 4
        ys, xs = np.divmod(np.arange(mb_w * mb_h), mb_w)
        for i in range(0, mb_w * mb_h, 97):
            x0, y0 = xs[i] * 16, ys[i] * 16
            blk = refpad[y0 + mv[i, 1] + 48:y0 + mv[i, 1] + 64, x0 + mv[i, 0] + 48:x0 + mv[i, 0] + 64]
            assert int(np.abs(src[y0:y0 + 16, x0:x0 + 16] - blk).sum()) == int(got[r]["sad"][i, 0])


def test_motion_search_rejects_empty_slot():
    with Analyser(64, 64, nrefs=2, search_mode=abi.SEARCH_FULL, search_range=8) as an:
        pics = frames(64, 64, 1, 3)
        an.set_reference(0, pics[0])
        an.upload(pics[0])
        with pytest.raises(Exception):
            an.motion_search([0, 1])
