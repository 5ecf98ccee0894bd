"""The workloads at their own sizes on the GPU.

- 1280x720 (the reference's CPU-runnable configuration): QP 28, 3 references, the reference's diamond
  search, 5 slices (min(cores, mb_h/8) on a host with 5 or more cores, enc/mod.rs:192), IDR + 3 P
  pictures, every picture bit-exact with the oracle.
- 1920x1080 learned proposal (+-32, 3 references, trained weights, fallback at 1.1x the predicted
  cost): every returned SAD exact at its vector, partitions sampled against numpy; the fallback rate
  and the best reference's mean 16x16 SAD gap against the exhaustive search at most 10% and 1%;
  with a fallback ratio of 0 the results are the exhaustive search's exactly.
- 64-stream mix, one stream of each size (640x360 ... 3840x2160, objects +-24 px/picture plus a pan of
  +-4, +-32 exhaustive search, 3 references): four frame steps replayed as one hipGraph, each stream
  equal to the single-context path, and the 1920x1080 stream's chain equal to the oracle (its
  references loop-filtered by the oracle).
- learned + batched at 3840x2160 (+-64, objects +-64, noise 4, a scene cut): IDR pictures where the
  encoder driver's scene-cut test puts them, two streams through the replayed graph equal to the
  single-context path, SADs exact at the learned vectors (sampled).
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen, h264w, proposal
from ironav_amd.h264a import Analyser, Batch, Picture

pytestmark = pytest.mark.gpu


def frames(spec):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


def compare(tag, got, want):
    gc, gs, grec, gm = got
    wc, ws, wrec, wm = want
    for f in abi.MBCODE_DTYPE.names:
        if not np.array_equal(gc[f], wc[f]):
            bad = np.nonzero(np.any((gc[f] != wc[f]).reshape(len(gc), -1), axis=1))[0]
            raise AssertionError(f"{tag}: code.{f} differs at {len(bad)} macroblocks, first {bad[0]}")
    for f in abi.MBSTATS_DTYPE.names:
        assert np.array_equal(gs[f], ws[f]), f"{tag}: stats.{f} differs"
    for k in range(3):
        assert np.array_equal(grec[k], wrec[k]), f"{tag}: reconstruction plane {k} differs"
    for f in abi.MBINFO_DTYPE.names:
        assert np.array_equal(gm[f], wm[f]), f"{tag}: mbinfo.{f} differs"


def same(a, b):
    return (np.array_equal(a.codes, b.codes) and np.array_equal(a.stats, b.stats) and np.array_equal(a.mbs, b.mbs)
            and np.array_equal(a.recon.y, b.recon.y) and np.array_equal(a.recon.cb, b.recon.cb)
            and np.array_equal(a.recon.cr, b.recon.cr))


def sad_at(src, refpad, x0, y0, w, h, dx, dy):
    blk = refpad[y0 + dy + 48:y0 + dy + 48 + h, x0 + dx + 48:x0 + dx + 48 + w]
    return int(np.abs(src[y0:y0 + h, x0:x0 + w].astype(np.int32) - blk).sum())


def check_sads_exact(src, refs, got, mb_w, sample, every=1):
    """Every 16x16 SAD (each `every`-th macroblock) and all 41 partitions of `sample` macroblocks equal
    the SAD recomputed at the returned vector."""
    s32 = src.astype(np.int32)
    for r, ref in enumerate(refs):
        pad = np.pad(ref.astype(np.int32), 48, mode="edge")
        mv = got[r]["mv"][:, 0] // 4
        for i in range(0, got.shape[1], every):
            x0, y0 = (i % mb_w) * 16, (i // mb_w) * 16
            blk = pad[y0 + mv[i, 1] + 48:y0 + mv[i, 1] + 64, x0 + mv[i, 0] + 48:x0 + mv[i, 0] + 64]
            assert int(np.abs(s32[y0:y0 + 16, x0:x0 + 16] - blk).sum()) == int(got[r]["sad"][i, 0]), (r, i)
        for i in sample:
            for p, (x, y, w, h) in enumerate(abi.PARTITIONS):
                dx, dy = (int(v) // 4 for v in got[r, i]["mv"][p])
                x0, y0 = (i % mb_w) * 16 + x * 4, (i // mb_w) * 16 + y * 4
                assert sad_at(src, pad, x0, y0, w * 4, h * 4, dx, dy) == int(got[r, i]["sad"][p]), (r, i, p)


def test_720p_chain_matches_oracle(oracle):
    """1280x720, QP 28 (IDR 25), 3 references, diamond search, 5 slices, CABAC + 8x8 transform."""
    pics = frames(clipgen.ClipSpec(1280, 720, 4, 720, max_motion=12, noise_sigma=2.0))
    with Analyser(1280, 720, nrefs=3, slices=5) as an:
        refs = []
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            q = 25 if t == 0 else 28
            nref = min(len(refs), 3) if t else 0
            for s in range(nref):
                an.set_reference(s, refs[s][0], refs[s][1])
            res = an.analyse(pic, q, kind, list(range(nref)))
            want = oracle.analyse_picture(an.cfg, pic, q, kind, refs[:nref], None, threads=8)
            compare(f"720p frame {t}", (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs), want)
            refs.insert(0, (res.recon, res.mbs))


def _references_1080p(pics, nref=3):
    """IDR + nref P pictures coded by the analysis (68 slices, diamond search), newest first, so that
    reference 0 carries the encoder's vectors."""
    with Analyser(1920, 1080, nrefs=3, slices=68) as an:
        res = an.analyse(pics[0], 25, abi.PIC_IDR)
        refs = [(res.recon, res.mbs)]
        for t in range(1, nref + 1):
            for s in range(min(len(refs), 3)):
                an.set_reference(s, refs[s][0], refs[s][1])
            res = an.analyse(pics[t], 28, abi.PIC_P, list(range(min(len(refs), 3))))
            refs.insert(0, (res.recon, res.mbs))
    return refs[:nref]


def _search(refs, pic, mode, ratio=1.1, params=None, rng=32):
    with Analyser(1920, 1080, nrefs=3, slices=68, search_mode=mode, search_range=rng, fallback_ratio=ratio) as an:
        if params is not None:
            an.load_proposal_weights(params)
        for s, (r, m) in enumerate(refs):
            an.set_reference(s, r, m)
        an.upload(pic)
        an.motion_search(list(range(len(refs))))
        got = an.read_motion(len(refs))
        fb = an.fallback_count() if mode == abi.SEARCH_LEARNED else 0
    return got, fb


def test_learned_proposal_1080p():
    pics = frames(clipgen.ClipSpec(1920, 1080, 5, 1080, max_motion=12, noise_sigma=2.0))  # the workload clip's first 5
    refs = _references_1080p(pics)
    params = proposal.load_weights()
    learned, fb = _search(refs, pics[4], abi.SEARCH_LEARNED, 1.1, params)
    full, _ = _search(refs, pics[4], abi.SEARCH_FULL)
    n = 120 * 68
    rs = np.random.default_rng(4)
    sample = sorted({0, 119, n - 120, n - 1, 60} | set(int(v) for v in rs.integers(0, n, 40)))
    check_sads_exact(pics[4].y, [r[0].y for r in refs], learned, 120, sample)
    rate = fb / (3 * n)
    best_l = learned["sad"][:, :, 0].min(axis=0).astype(np.float64)
    best_f = full["sad"][:, :, 0].min(axis=0).astype(np.float64)
    gap = best_l.mean() / best_f.mean() - 1.0
    print(f"learned 1080p: fallback rate {rate:.4f}, best-reference 16x16 SAD gap {gap:+.4%}")
    assert rate <= 0.10 and gap <= 0.01
    # with a fallback ratio of 0 every pair is searched exhaustively: the exhaustive search's results
    allfb, fb0 = _search(refs, pics[4], abi.SEARCH_LEARNED, 0.0, params)
    assert fb0 == 3 * n
    assert np.array_equal(allfb["sad"], full["sad"]) and np.array_equal(allfb["mv"], full["mv"])


SIZES = [(640, 360), (854, 480), (1280, 720), (1920, 1080), (3840, 2160)]
NFR = 4


def _mixed_streams():
    specs = [clipgen.ClipSpec(w, h, NFR, 3000 + i, max_motion=24, pan=4) for i, (w, h) in enumerate(SIZES)]
    slices = [max(1, ((h + 15) // 16) // 2) for (w, h) in SIZES]
    return specs, slices


def _schedule(t):
    refs = [(t - 1 - k) % 3 for k in range(min(t, 3))]
    return (25 if t == 0 else 28), (abi.PIC_IDR if t == 0 else abi.PIC_P), refs, t % 3


def test_mixed_size_batch_graph(oracle):
    specs, slices = _mixed_streams()
    clips = [frames(s) for s in specs]

    def make(i):
        w, h = SIZES[i]
        return Analyser(w, h, nrefs=3, slices=slices[i], search_mode=abi.SEARCH_FULL, search_range=32)

    # the single-context path, and the 1920x1080 stream's chain against the oracle
    want = []
    k1080 = SIZES.index((1920, 1080))
    for i in range(len(SIZES)):
        res = []
        refs_host = {}
        with make(i) as an:
            for t in range(NFR):
                qp, kind, refs, promote = _schedule(t)
                fs = None
                if i == k1080 and kind == abi.PIC_P:
                    an.upload(clips[i][t])
                    an.motion_search(refs)
                    fs = an.read_motion(len(refs))
                r = an.analyse(clips[i][t], qp, kind, refs)
                if i == k1080:
                    w_ = oracle.analyse_picture(an.cfg, clips[i][t], qp, kind, [refs_host[s] for s in refs], fs,
                                                threads=8)
                    compare(f"1080p stream frame {t}", (r.codes, r.stats, (r.recon.y, r.recon.cb, r.recon.cr), r.mbs), w_)
                    refs_host[promote] = (Picture(*oracle.deblock_picture(w_[2][0], w_[2][1], w_[2][2], w_[3],
                                                                          an.mb_w, an.mb_h)), w_[3])
                an.promote_reference(promote)
                res.append(r)
        want.append(res)
    # one hipGraph per frame step, replayed
    ans = [make(i) for i in range(len(SIZES))]
    with Batch(ans, use_graph=True) as b:
        for t in range(NFR):
            frs = []
            for i, an in enumerate(ans):
                qp, kind, refs, promote = _schedule(t)
                an.upload(clips[i][t])
                frs.append(dict(qp=qp, kind=kind, refs=refs, promote=promote))
            b.step(frs)
            b.sync()
            for i, an in enumerate(ans):
                assert same(an.read_analysis(), want[i][t]), (SIZES[i], t)
    for an in ans:
        an.close()


def test_learned_batched_2160p_with_scene_cut():
    w, h, nst, nfr = 3840, 2160, 2, 4
    specs = [clipgen.ClipSpec(w, h, nfr, 8000 + i, max_motion=64, noise_sigma=4.0, cuts=(2 + i,)) for i in range(nst)]
    clips = [frames(s) for s in specs]
    # picture types from the encoder driver's scene-cut test
    kinds = []
    for i in range(nst):
        with h264w.Gop(w, h) as g:
            kinds.append([g.next(p.y) for p in clips[i]])
    assert kinds == [[abi.PIC_IDR, abi.PIC_P, abi.PIC_IDR, abi.PIC_P], [abi.PIC_IDR, abi.PIC_P, abi.PIC_P, abi.PIC_IDR]]
    params = proposal.load_weights()

    def sched(i, t):
        last = max(u for u in range(t + 1) if kinds[i][u] == abi.PIC_IDR)
        n = min(t - last, 3)
        refs = [(t - 1 - k) % 3 for k in range(n)]
        kind = kinds[i][t]
        return (25 if kind == abi.PIC_IDR else 28), kind, refs, t % 3

    def make():
        an = Analyser(w, h, nrefs=3, slices=68, search_mode=abi.SEARCH_LEARNED, search_range=64, fallback_ratio=1.1)
        an.load_proposal_weights(params)
        return an

    want = []
    for i in range(nst):
        res = []
        with make() as an:
            for t in range(nfr):
                qp, kind, refs, promote = sched(i, t)
                r = an.analyse(clips[i][t], qp, kind, refs)
                if kind == abi.PIC_P:
                    got = an.read_motion(len(refs))
                    refy = [an.read_reference(s)[0].y for s in refs]
                    rs = np.random.default_rng(t)
                    sample = sorted(set(int(v) for v in rs.integers(0, 240 * 135, 12)))
                    check_sads_exact(clips[i][t].y, refy, got, 240, sample, every=37)
                an.promote_reference(promote)
                res.append(r)
        want.append(res)
    ans = [make() for _ in range(nst)]
    with Batch(ans, use_graph=True) as b:
        for t in range(nfr):
            frs = []
            for i, an in enumerate(ans):
                qp, kind, refs, promote = sched(i, t)
                an.upload(clips[i][t])
                frs.append(dict(qp=qp, kind=kind, refs=refs, promote=promote))
            b.step(frs)
            b.sync()
            for i, an in enumerate(ans):
                assert same(an.read_analysis(), want[i][t]), (i, t)
    for an in ans:
        an.close()
