"""The batched path on the GPU: streams of different sizes and settings through one set of launches
per frame step (eager and as a replayed hipGraph) give, per stream and per frame, exactly what the
single-context path gives (itself bit-exact with the oracle), including the on-device promotion of
each reconstruction to a reference, loop-filtered (as the encoder keeps its references) or not."""
import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Analyser, Batch, Picture

pytestmark = pytest.mark.gpu

# (width, height, seed, search range, nrefs, slices, cabac, transform8)
STREAMS = [(64, 48, 1, 16, 2, 1, 1, 1), (96, 64, 2, 16, 2, 2, 1, 1), (160, 96, 3, 32, 3, 3, 1, 0),
           (48, 32, 4, 8, 1, 2, 0, 0), (128, 128, 5, 16, 2, 4, 1, 1)]
FRAMES = 4


def clip(w, h, seed):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, FRAMES, seed))]


def schedule(nrefs, t):
    """IDR at 0, then P pictures against the newest min(t, nrefs) references; each picture becomes a
    reference in slot t % nrefs."""
    refs = [(t - 1 - k) % nrefs for k in range(min(t, nrefs))]
    return (25 if t == 0 else 28 + (t % 3) - 1), (abi.PIC_IDR if t == 0 else abi.PIC_P), refs, t % nrefs


def analyser(w, h, rng, nrefs, slices, cabac, t8, fast=False):
    # the live mode searches as the reference's does (diamond); the default streams add the exhaustive winner
    return Analyser(w, h, nrefs=nrefs, slices=slices, cabac=bool(cabac), transform8=bool(t8),
                    search_mode=abi.SEARCH_DIAMOND if fast else abi.SEARCH_FULL, search_range=rng, fast=fast)


def single_runs(deblock=True, fast=False):
    out = []
    for (w, h, seed, rng, nrefs, slices, cabac, t8) in STREAMS:
        pics = clip(w, h, seed)
        res = []
        with analyser(w, h, rng, nrefs, slices, cabac, t8, fast) as an:
            for t in range(FRAMES):
                qp, kind, refs, promote = schedule(nrefs, t)
                r = an.analyse(pics[t], qp, kind, refs)
                an.promote_reference(promote, deblock=deblock)
                res.append(r)
        out.append(res)
    return out


def batch_runs(use_graph, device_planes=False, deblock=True, fast=False):
    import torch

    clips = [clip(w, h, seed) for (w, h, seed, *_rest) in STREAMS]
    ans = [analyser(w, h, rng, nrefs, slices, cabac, t8, fast)
           for (w, h, seed, rng, nrefs, slices, cabac, t8) in STREAMS]
    out = [[] for _ in STREAMS]
    keep = []
    with Batch(ans, use_graph=use_graph) as b:
        for t in range(FRAMES):
            frames = []
            for i, an in enumerate(ans):
                qp, kind, refs, promote = schedule(STREAMS[i][4], t)
                d = dict(qp=qp, kind=kind, refs=refs, promote=promote, deblock=deblock)
                pic = clips[i][t]
                if device_planes:
                    planes = [torch.from_numpy(p).cuda() for p in (pic.y, pic.cb, pic.cr)]
                    keep.append(planes)
                    d["planes"] = tuple(p.data_ptr() for p in planes)
                else:
                    an.upload(pic)
                frames.append(d)
            torch.cuda.synchronize()
            b.step(frames)
            b.sync()
            for i, an in enumerate(ans):
                out[i].append(an.read_analysis())
    for an in ans:
        an.close()
    return out


def same(a, b):
    return (np.array_equal(a.codes, b.codes) and np.array_equal(a.stats, b.stats) and np.array_equal(a.mbs, b.mbs)
            and np.array_equal(a.recon.y, b.recon.y) and np.array_equal(a.recon.cb, b.recon.cb)
            and np.array_equal(a.recon.cr, b.recon.cr))


@pytest.fixture(scope="module")
def single():
    return single_runs()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_streams(single, use_graph):
    got = batch_runs(use_graph)
    for i in range(len(STREAMS)):
        for t in range(FRAMES):
            assert same(got[i][t], single[i][t]), (STREAMS[i], t)


def test_batch_graph_from_device_planes(single):
    """Pictures handed as device planes (no upload); the graph replayed every frame."""
    got = batch_runs(True, device_planes=True)
    for i in range(len(STREAMS)):
        for t in range(FRAMES):
            assert same(got[i][t], single[i][t]), (STREAMS[i], t)


def test_batch_without_loop_filter():
    """Promotion without the loop filter (frames' promote_deblock 0) against the single-context
    promotion without it."""
    want = single_runs(deblock=False)
    got = batch_runs(True, deblock=False)
    for i in range(len(STREAMS)):
        for t in range(FRAMES):
            assert same(got[i][t], want[i][t]), (STREAMS[i], t)


def test_batch_live_mode_streams():
    """Streams in the live mode (realtime decisions, diamond search; CABAC and CAVLC), replayed as a
    graph, against the single-context path in the same mode."""
    want = single_runs(fast=True)
    got = batch_runs(True, fast=True)
    for i in range(len(STREAMS)):
        for t in range(FRAMES):
            assert same(got[i][t], want[i][t]), (STREAMS[i], t)


def test_single_stream_chain_matches_oracle(oracle):
    """The single-context chain the batch is compared with, pinned to the oracle for one stream:
    every reference loop-filtered (the oracle's filter of the oracle's reconstruction)."""
    w, h, seed, rng, nrefs, slices, cabac, t8 = STREAMS[1]
    pics = clip(w, h, seed)
    with analyser(w, h, rng, nrefs, slices, cabac, t8) as an:
        refs_host = {}
        for t in range(3):
            qp, kind, refs, promote = schedule(nrefs, t)
            if kind == abi.PIC_P:
                an.upload(pics[t])
                an.motion_search(refs)
                fs = an.read_motion(len(refs))
            r = an.analyse(pics[t], qp, kind, refs)
            want = oracle.analyse_picture(an.cfg, pics[t], qp, kind, [refs_host[s] for s in refs],
                                          fs if kind == abi.PIC_P else None)
            assert np.array_equal(r.codes, want[0]) and np.array_equal(r.stats, want[1]), t
            assert np.array_equal(r.recon.y, want[2][0]) and np.array_equal(r.mbs, want[3]), t
            an.promote_reference(promote)
            refs_host[promote] = (Picture(*oracle.deblock_picture(want[2][0], want[2][1], want[2][2], want[3],
                                                                  an.mb_w, an.mb_h)), want[3])


# learned streams (trained weights), one with every pair falling back, beside a full-search stream
LSTREAMS = [(96, 64, 21, 16, 3, 2, 1.1), (160, 96, 22, 32, 2, 3, 1.1), (64, 48, 23, 16, 2, 1, 0.0),
            (128, 96, 24, 32, 3, 2, float("inf"))]


def learned_analyser(w, h, rng, nrefs, slices, ratio, params):
    an = Analyser(w, h, nrefs=nrefs, slices=slices, search_mode=abi.SEARCH_LEARNED, search_range=rng,
                  fallback_ratio=ratio)
    an.load_proposal_weights(params)
    return an


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_with_learned_streams_matches_single_streams(use_graph):
    from ironav_amd import proposal

    params = proposal.load_weights()
    clips = [clip(w, h, seed) for (w, h, seed, *_r) in LSTREAMS] + [clip(*STREAMS[1][:3])]
    # single-context runs
    want = []
    for i, spec in enumerate(LSTREAMS + [STREAMS[1]]):
        an = learned_analyser(*spec[:2], spec[3], spec[4], spec[5], spec[6], params) if i < len(LSTREAMS) \
            else analyser(*spec[:2], *spec[3:])
        res = []
        with an:
            for t in range(FRAMES):
                qp, kind, refs, promote = schedule(an.cfg.nrefs, t)
                r = an.analyse(clips[i][t], qp, kind, refs)
                an.promote_reference(promote)
                res.append(r)
        want.append(res)
    # the batch
    ans = [learned_analyser(*s[:2], s[3], s[4], s[5], s[6], params) for s in LSTREAMS] + [analyser(*STREAMS[1][:2], *STREAMS[1][3:])]
    with Batch(ans, use_graph=use_graph) as b:
        for t in range(FRAMES):
            frames = []
            for i, an in enumerate(ans):
                qp, kind, refs, promote = schedule(an.cfg.nrefs, t)
                an.upload(clips[i][t])
                frames.append(dict(qp=qp, kind=kind, refs=refs, promote=promote))
            b.step(frames)
            b.sync()
            nfb = b.fallback_count()
            if t == 0:
                assert nfb == 0
            else:
                # the stream with ratio 0 sends every (macroblock, reference)
                assert nfb >= 12 * min(t, 2)
            for i, an in enumerate(ans):
                assert same(an.read_analysis(), want[i][t]), (i, t)
    for an in ans:
        an.close()


# B pictures in the batch: CABAC streams (width, height, seed, search range, nrefs, slices) coded with
# two B pictures between references, beside a CAVLC stream coding IDR and P pictures in the same steps
B_STREAMS = [(64, 48, 11, 16, 2, 1), (96, 64, 12, 16, 3, 2), (128, 96, 13, 32, 2, 3)]
B_ORDER = [0, 3, 1, 2, 6, 4, 5]  # display index of each step's picture (the encoder's coding order)


def b_schedule(nrefs, t):
    """Step t's frame: references (display index a multiple of 3) go to slot (index / 3) % nrefs and
    are predicted from the references before them, newest first; a B picture lies between the
    references before and after it."""
    d = B_ORDER[t]
    k = d // 3
    if d % 3 == 0:
        if d == 0:
            return dict(qp=25, kind=abi.PIC_IDR, refs=[], promote=0)
        return dict(qp=28, kind=abi.PIC_P, refs=[(k - 1 - j) % nrefs for j in range(min(k, nrefs))], promote=k % nrefs)
    return dict(qp=31, kind=abi.PIC_B, refs=[k % nrefs, (k + 1) % nrefs], promote=-1, dist=(d - 3 * k, 3))


def b_clip(w, h, seed):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, len(B_ORDER), seed))]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_b_pictures_match_single_streams(use_graph):
    """B pictures through the batched step (eager and as a replayed graph) equal the single-context
    path's ia_analyse_b, with the P and IDR pictures of the same streams and a CAVLC stream's P
    pictures in the same steps."""
    import torch

    p_stream = STREAMS[3]  # CAVLC
    want = []
    for (w, h, seed, rng, nrefs, slices) in B_STREAMS:
        pics = b_clip(w, h, seed)
        res = []
        with analyser(w, h, rng, nrefs, slices, 1, 1) as an:
            for t in range(len(B_ORDER)):
                f = b_schedule(nrefs, t)
                pic = pics[B_ORDER[t]]
                if f["kind"] == abi.PIC_B:
                    res.append(an.analyse_b(pic, f["qp"], f["refs"][0], f["refs"][1], *f["dist"]))
                else:
                    res.append(an.analyse(pic, f["qp"], f["kind"], f["refs"]))
                    an.promote_reference(f["promote"], deblock=True)
        want.append(res)
    (pw, ph, pseed, prng, pnrefs, pslices, pcabac, pt8) = p_stream
    ppics = b_clip(pw, ph, pseed)
    pwant = []
    with analyser(pw, ph, prng, pnrefs, pslices, pcabac, pt8) as an:
        for t in range(len(B_ORDER)):
            qp, kind, refs, promote = schedule(pnrefs, t)
            pwant.append(an.analyse(ppics[t], qp, kind, refs))
            an.promote_reference(promote, deblock=True)

    clips = [b_clip(w, h, seed) for (w, h, seed, *_rest) in B_STREAMS]
    ans = [analyser(w, h, rng, nrefs, slices, 1, 1) for (w, h, seed, rng, nrefs, slices) in B_STREAMS]
    ans.append(analyser(pw, ph, prng, pnrefs, pslices, pcabac, pt8))
    try:
        with Batch(ans, use_graph=use_graph) as b:
            for t in range(len(B_ORDER)):
                frames = []
                for i, (w, h, seed, rng, nrefs, slices) in enumerate(B_STREAMS):
                    ans[i].upload(clips[i][B_ORDER[t]])
                    frames.append(dict(b_schedule(nrefs, t), deblock=True))
                qp, kind, refs, promote = schedule(pnrefs, t)
                ans[-1].upload(ppics[t])
                frames.append(dict(qp=qp, kind=kind, refs=refs, promote=promote, deblock=True))
                torch.cuda.synchronize()
                b.step(frames)
                b.sync()
                for i in range(len(B_STREAMS)):
                    assert same(ans[i].read_analysis(), want[i][t]), (B_STREAMS[i], t, B_ORDER[t])
                assert same(ans[-1].read_analysis(), pwant[t]), (p_stream, t)
    finally:
        for an in ans:
            an.close()


def test_batch_refuses_b_picture_as_reference():
    with analyser(64, 48, 16, 2, 1, 1, 1) as an, Batch([an], use_graph=False) as b:
        pic = b_clip(64, 48, 3)[0]
        an.upload(pic)
        b.step([dict(qp=25, kind=abi.PIC_IDR, refs=[], promote=0)])
        b.sync()
        an.upload(pic)
        b.step([dict(qp=28, kind=abi.PIC_P, refs=[0], promote=1)])
        b.sync()
        an.upload(pic)
        with pytest.raises(ValueError, match="not a reference"):
            b.step([dict(qp=31, kind=abi.PIC_B, refs=[0, 1], promote=0, dist=(1, 2))])
        with pytest.raises(ValueError, match="0 < dist_l0 < dist_l1"):
            b.step([dict(qp=31, kind=abi.PIC_B, refs=[0, 1], promote=-1, dist=(2, 2))])
