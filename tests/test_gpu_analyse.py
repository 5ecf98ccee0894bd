"""GPU parity of the per-macroblock analysis (code_mb over every slice) against the oracle.

Bit-exact on the same inputs: every field of every macroblock's code (kind, skip, modes, partition,
vectors and their differences, references, sub-partitions, levels, coded-block pattern, coefficient
counts), the costs the decisions were taken on (P_Skip SAD, Intra16x16 cost, best inter cost and
partitioning, and the rate-distortion J as an IEEE double), the reconstruction, and the MbInfo
records. Pictures chain IPPP: each P picture's references are earlier reconstructions, most recent
first (the deblocking between pictures is the caller's and is left out: both sides get the same
undeblocked reference).
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Analyser, Picture

pytestmark = pytest.mark.gpu


def frames(w, h, n, seed, **kw):
    spec = clipgen.ClipSpec(w, h, n, seed, **kw)
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


def compare(tag, got, want):
    gc, gs, grec, gm = got
    wc, ws, wrec, wm = want
    for f in abi.MBCODE_DTYPE.names:
        if not np.array_equal(gc[f], wc[f]):
            bad = np.nonzero(np.any((gc[f] != wc[f]).reshape(len(gc), -1), axis=1))[0]
            i = int(bad[0])
            raise AssertionError(f"{tag}: code.{f} differs at {len(bad)} macroblocks, first {i}: "
                                 f"kind {gc['kind'][i]}/{wc['kind'][i]} got {gc[f][i]} want {wc[f][i]}")
    for f in abi.MBSTATS_DTYPE.names:
        if not np.array_equal(gs[f], ws[f]):
            bad = np.nonzero(gs[f] != ws[f])[0]
            raise AssertionError(f"{tag}: stats.{f} differs at {len(bad)} macroblocks, first {bad[0]}: "
                                 f"{gs[f][bad[0]]} vs {ws[f][bad[0]]}")
    for k, name in enumerate("y cb cr".split()):
        assert np.array_equal(grec[k], wrec[k]), f"{tag}: reconstruction {name} differs"
    for f in abi.MBINFO_DTYPE.names:
        assert np.array_equal(gm[f], wm[f]), f"{tag}: mbinfo.{f} differs"


def run_chain(oracle, w, h, nframes, qp, *, nrefs=3, slices=1, cabac=True, transform8=True, sub8x8=False,
              search_mode=abi.SEARCH_DIAMOND, search_range=16, seed=1, fast=False, **clip):
    pics = frames(w, h, nframes, seed, **clip)
    cfg = dict(nrefs=nrefs, slices=slices, cabac=cabac, transform8=transform8, sub8x8=sub8x8,
               search_mode=search_mode, search_range=search_range, fast=fast)
    with Analyser(w, h, **cfg) as an:
        refs = []  # (Picture, mbinfo), most recent first
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            q = qp - 3 if kind == abi.PIC_IDR else qp
            nref = min(len(refs), nrefs)
            slots = list(range(nref))
            for s in slots:
                an.set_reference(s, refs[s][0], refs[s][1])
            fs = None
            if kind == abi.PIC_P and search_mode == abi.SEARCH_FULL:
                an.upload(pic)
                an.motion_search(slots)
                fs = an.read_motion(nref)
                want_fs = oracle.fullsearch(pic.y, [r[0].y for r in refs[:nref]], an.mb_w, an.mb_h, search_range)
                assert np.array_equal(fs, want_fs)
            res = an.analyse(pic, q, kind, slots)
            got = (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs)
            want = oracle.analyse_picture(an.cfg, pic, q, kind, refs[:nref], fs)
            compare(f"frame {t} qp {q}", got, want)
            refs.insert(0, (res.recon, res.mbs))
    return got


@pytest.mark.parametrize("qp", [10, 28, 45])
def test_idr_picture(oracle, qp):
    pics = frames(96, 64, 1, 5)
    with Analyser(96, 64, nrefs=1, slices=2) as an:
        res = an.analyse(pics[0], qp, abi.PIC_IDR)
        want = oracle.analyse_picture(an.cfg, pics[0], qp, abi.PIC_IDR)
    compare("idr", (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs), want)


@pytest.mark.parametrize("qp,nrefs,slices", [(28, 1, 1), (22, 3, 2), (34, 2, 3)])
def test_ippp_chain_diamond(oracle, qp, nrefs, slices):
    run_chain(oracle, 128, 96, 4, qp, nrefs=nrefs, slices=slices, seed=qp)


def test_ippp_chain_full_search(oracle):
    run_chain(oracle, 128, 96, 4, 28, nrefs=3, slices=2, search_mode=abi.SEARCH_FULL, search_range=24, seed=9,
              max_motion=10)


def test_sub8x8_at_low_qp(oracle):
    run_chain(oracle, 96, 64, 3, 18, nrefs=2, slices=1, sub8x8=True, seed=4)


def test_cavlc_satd_decisions(oracle):
    run_chain(oracle, 96, 64, 3, 26, nrefs=2, slices=2, cabac=False, transform8=False, seed=6)


def test_no_8x8_transform(oracle):
    run_chain(oracle, 80, 48, 3, 30, nrefs=1, slices=1, transform8=False, seed=8)


def test_odd_sizes_and_single_macroblock(oracle):
    run_chain(oracle, 16, 16, 3, 28, nrefs=2, slices=1, seed=2)
    run_chain(oracle, 40, 26, 3, 24, nrefs=1, slices=2, seed=3)


def test_extreme_qps(oracle):
    run_chain(oracle, 64, 48, 3, 51, nrefs=1, slices=1, seed=12)
    run_chain(oracle, 64, 48, 3, 3, nrefs=1, slices=1, seed=13)


def test_large_motion_and_noise(oracle):
    run_chain(oracle, 160, 96, 3, 28, nrefs=2, slices=2, seed=21, max_motion=24, noise_sigma=4.0)


@pytest.mark.parametrize("qp", [22, 28, 34])
def test_1080p_workload_chain(oracle, qp):
    """The plain GPU path at the workload's size: 1920x1080, IDR + 2 P pictures, 3 references, 8
    slices, exhaustive search +-32 as the extra candidate, 8x8 transform, CABAC; bit-exact."""
    run_chain(oracle, 1920, 1080, 3, qp, nrefs=3, slices=8, search_mode=abi.SEARCH_FULL, search_range=32,
              seed=1080 + qp)


@pytest.mark.parametrize("cabac,qp", [(True, 26), (False, 26), (False, 38)])
def test_realtime_decisions(oracle, cabac, qp):
    """The live mode's decisions (enc/frame.rs:731-765, 861-863, 973, 990-992, 1006): early P_Skip
    without a transform, 16x16 only, the shorter diamond, the integer stop, half samples only, and
    Intra16x16 alone for flat macroblocks of CAVLC I pictures."""
    run_chain(oracle, 128, 96, 4, qp, nrefs=1, slices=2, cabac=cabac, transform8=cabac, fast=True, seed=51 + qp,
              max_motion=8)


def test_realtime_still_scene_and_full_search(oracle):
    """A still scene takes the untransformed P_Skip in most macroblocks; the exhaustive winners as the
    extra candidate of the shorter search."""
    got = run_chain(oracle, 96, 64, 3, 30, nrefs=1, slices=1, fast=True, seed=60, max_motion=0, noise_sigma=0.0)
    assert (got[0]["skip"] == 1).mean() > 0.5
    run_chain(oracle, 128, 96, 3, 22, nrefs=2, slices=2, fast=True, search_mode=abi.SEARCH_FULL, search_range=24,
              seed=61, max_motion=10)


def test_realtime_1080p(oracle):
    """The live mode at 1920x1080: IDR + 2 P pictures, one reference, 34 slices (one per two
    macroblock rows, as `realtime()` sets them on a many-core host), CABAC; bit-exact."""
    from ironav_amd.h264a import realtime
    kw = realtime(1080, 64)
    run_chain(oracle, 1920, 1080, 3, 28, nrefs=kw["nrefs"], slices=kw["slices"], fast=True, seed=1081)
