"""The learned motion proposal on the GPU: the BF16 MFMA CNN against an FP32 reference, the FP32
training kernels against torch autograd on the CPU, and the learned search's invariants."""
import ctypes as C

import numpy as np
import pytest

from ironav_amd import abi, clipgen, proposal
from ironav_amd.h264a import Analyser, Picture

pytestmark = pytest.mark.gpu


def frames(w, h, n, seed, **kw):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, n, seed, **kw))]


def test_cnn_forward_bf16_mfma_matches_fp32_reference():
    """Tolerance: BF16 operands (8-bit significands) at every layer with FP32 accumulation; outputs
    within 3% of the output range of the FP32 network."""
    rng = np.random.default_rng(0)
    params = proposal.init_params(3)
    feat = rng.uniform(0.0, 1.0, size=(24, 81, 8)).astype(np.float32)
    feat[:, :, 3:6] = (rng.random((24, 81, 3)) < 0.05).astype(np.float32)
    feat[:, :, 6] = 1.0
    with Analyser(96, 64, nrefs=1) as an:
        an.load_proposal_weights(params)
        got = an.proposal_forward(feat)
    want = proposal.forward_reference(params, feat)
    scale = np.abs(want).max()
    assert scale > 0.1
    err = np.abs(got - want).max()
    assert err <= 0.03 * scale, (err, scale)
    # the argmax cells mostly agree (ties aside)
    agree = np.mean(np.argmax(got[:, :, :3], axis=1) == np.argmax(want[:, :, :3], axis=1))
    assert agree > 0.9


def test_training_kernels_match_torch_autograd():
    """Gradients of the HIP FP32 training kernels against torch autograd of the same network and loss
    on the CPU (torch fp32 reference); tolerance 1e-3 relative to the largest gradient."""
    import torch

    rng = np.random.default_rng(1)
    B = 8
    params = proposal.init_params(5)
    feat = rng.uniform(0, 1, size=(B, 81, 8)).astype(np.float32)
    cell = rng.integers(-1, 81, size=(B, 3)).astype(np.int32)
    cost = rng.uniform(4, 12, size=(B, 3)).astype(np.float32)
    cost_w = 0.05
    # torch reference (CPU, fp32)
    P = torch.tensor(params, requires_grad=True)
    x = torch.tensor(feat).reshape(B, 9, 9, 8).permute(0, 3, 1, 2)
    for li, ((w, b), (cin, cout)) in enumerate(zip(proposal.param_slices(), proposal.LAYERS)):
        x = torch.nn.functional.conv2d(x, P[w[0]:w[1]].reshape(cout, cin, 3, 3), P[b[0]:b[1]], padding=1)
        if li < 3:
            x = torch.relu(x)
    out = x.permute(0, 2, 3, 1).reshape(B, 81, 4)
    loss = torch.zeros(())
    for n in range(B):
        for r in range(3):
            tc = int(cell[n, r])
            if tc < 0:
                continue
            loss = loss + (torch.logsumexp(out[n, :, r], 0) - out[n, tc, r] + cost_w * (out[n, tc, 3] - float(cost[n, r])) ** 2) / B
    loss.backward()
    want = P.grad.numpy()
    # HIP kernels: one batch, no update (lr 0)
    L = proposal._TrainLib()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    Pd = torch.tensor(params, device=dev)
    G = torch.zeros_like(Pd)
    X = torch.tensor(feat, device=dev)
    TC = torch.tensor(cell, device=dev)
    TK = torch.tensor(cost, device=dev)
    acts = [torch.empty((B, 81, co), device=dev) for _, co in proposal.LAYERS]
    grads = [torch.empty((B, 81, co), device=dev) for _, co in proposal.LAYERS]
    lossd = torch.zeros(1, device=dev)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    sl = proposal.param_slices()
    inp = X
    for li, ((w, b), (cin, cout)) in enumerate(zip(sl, proposal.LAYERS)):
        L.it_conv_fwd(ptr(inp), ptr(Pd[w[0]:]), ptr(Pd[b[0]:]), ptr(acts[li]), B, cin, cout, int(li < 3), stream)
        inp = acts[li]
    grads[3].zero_()
    L.it_loss_grad(ptr(acts[3]), ptr(TC), ptr(TK), ptr(grads[3]), ptr(lossd), B, cost_w, stream)
    for li in range(3, -1, -1):
        (w, b), (cin, cout) = sl[li], proposal.LAYERS[li]
        src = X if li == 0 else acts[li - 1]
        L.it_conv_bwd_w(ptr(src), ptr(grads[li]), ptr(G[w[0]:]), ptr(G[b[0]:]), B, cin, cout, stream)
        if li > 0:
            L.it_conv_bwd_data(ptr(grads[li]), ptr(Pd[w[0]:]), ptr(grads[li - 1]), B, cin, cout, stream)
            L.it_relu_bwd(ptr(grads[li - 1]), ptr(acts[li - 1]), B * 81 * cin, stream)
    torch.cuda.synchronize()
    got = G.cpu().numpy()
    assert abs(float(lossd.item()) - float(loss.item())) <= 1e-4 * max(1.0, abs(float(loss.item())))
    assert np.allclose(acts[3].cpu().numpy(), out.detach().numpy(), atol=1e-4, rtol=1e-4)
    assert np.abs(got - want).max() <= 1e-3 * np.abs(want).max(), np.abs(got - want).max()


def _learned(w, h, nref, ratio, params, seed=7, search_range=32, **clip):
    pics = frames(w, h, nref + 2, seed, **clip)
    with Analyser(w, h, nrefs=3, slices=2, search_mode=abi.SEARCH_LEARNED, search_range=search_range,
                  fallback_ratio=ratio) as an:
        an.load_proposal_weights(params)
        # reference 0 carries the encoder's vectors: analyse the pictures before
        res = an.analyse(pics[0], 25, abi.PIC_IDR)
        refs = [(res.recon, res.mbs)]
        for t in range(1, nref + 1):
            for s in range(len(refs)):
                an.set_reference(s, refs[s][0], refs[s][1])
            res = an.analyse(pics[t], 28, abi.PIC_P, list(range(len(refs))))
            refs.insert(0, (res.recon, res.mbs))
        for s in range(nref):
            an.set_reference(s, refs[s][0], refs[s][1])
        an.upload(pics[nref + 1])
        an.motion_search(list(range(nref)))
        got = an.read_motion(nref)
        fb = an.fallback_count()
    return pics[nref + 1], [r[0] for r in refs[:nref]], got, fb


def test_learned_search_sads_are_exact():
    params = proposal.init_params(11)
    src, refs, got, fb = _learned(128, 96, 3, 1e9, params, max_motion=8)
    assert fb == 0
    mb_w = 8
    for r in range(3):
        pad = np.pad(refs[r].y.astype(np.int32), 48, mode="edge")
        for mb in range(0, 48, 5):
            mbx, mby = mb % mb_w, mb // mb_w
            for p, (x, y, w, h) in enumerate(abi.PARTITIONS):
                dx, dy = (int(v) // 4 for v in got[r, mb]["mv"][p])
                x0, y0 = mbx * 16 + x * 4, mby * 16 + y * 4
                blk = pad[y0 + dy + 48:y0 + dy + 48 + h * 4, x0 + dx + 48:x0 + dx + 48 + w * 4]
                sad = int(np.abs(src.y[y0:y0 + h * 4, x0:x0 + w * 4].astype(np.int32) - blk).sum())
                assert sad == int(got[r, mb]["sad"][p]), (r, mb, p)
                assert -(mbx * 16 + 24) <= dx <= 128 - mbx * 16 + 8 and -(mby * 16 + 24) <= dy <= 96 - mby * 16 + 8


@pytest.mark.parametrize("w,h,rng", [(96, 64, 32), (208, 160, 32), (288, 224, 64)])
def test_learned_search_falls_back_to_full_search(oracle, w, h, rng):
    """With a fallback ratio of 0 every (macroblock, reference) goes to the exhaustive search through
    the fallback list: the results are the oracle's exhaustive search exactly (clipped windows at the
    small size; whole +-32 windows, 65 rows, inside the larger pictures; +-64 windows)."""
    params = proposal.init_params(12)
    src, refs, got, fb = _learned(w, h, 2, 0.0, params, search_range=rng, max_motion=rng // 2)
    mb_w, mb_h = w // 16, h // 16
    assert fb == mb_w * mb_h * 2
    want = oracle.fullsearch(src.y, [r.y for r in refs], mb_w, mb_h, rng)
    assert np.array_equal(got["sad"], want["sad"]) and np.array_equal(got["mv"], want["mv"])


def test_learned_analysis_runs_end_to_end():
    """IA_SEARCH_LEARNED through ia_analyse_picture: the learned winners are the extra candidates."""
    params = proposal.init_params(13)
    pics = frames(96, 64, 3, 8)
    with Analyser(96, 64, nrefs=2, slices=2, search_mode=abi.SEARCH_LEARNED, search_range=32) as an:
        an.load_proposal_weights(params)
        res = an.analyse(pics[0], 25, abi.PIC_IDR)
        an.set_reference(0, res.recon, res.mbs)
        p = an.analyse(pics[1], 28, abi.PIC_P, [0])
    assert np.all(p.codes["kind"] > 0)


def test_fallback_count_per_call_and_fallback_off(oracle):
    """The fallback count is per call (each call's list search clears the count the next call uses),
    and an infinite ratio turns the fallback off (no entries, the proposal's own results)."""
    params = proposal.init_params(14)
    pics = frames(96, 64, 3, 9)
    with Analyser(96, 64, nrefs=2, search_mode=abi.SEARCH_LEARNED, search_range=16, fallback_ratio=0.0) as an:
        an.load_proposal_weights(params)
        for s in range(2):
            an.set_reference(s, pics[s])
        an.upload(pics[2])
        want = oracle.fullsearch(pics[2].y, [pics[0].y, pics[1].y], 6, 4, 16)
        for nref in (2, 1, 2):
            an.motion_search(list(range(nref)))
            got = an.read_motion(nref)
            assert an.fallback_count() == 24 * nref
            assert np.array_equal(got["sad"], want["sad"][:nref]) and np.array_equal(got["mv"], want["mv"][:nref])
    with Analyser(96, 64, nrefs=2, search_mode=abi.SEARCH_LEARNED, search_range=16,
                  fallback_ratio=float("inf")) as an:
        an.load_proposal_weights(params)
        for s in range(2):
            an.set_reference(s, pics[s])
        an.upload(pics[2])
        for _ in range(2):
            an.motion_search([0, 1])
            assert an.fallback_count() == 0
        got = an.read_motion(2)
        # the proposal's own winners (its grid is not bounded by the exhaustive search's range)
        assert np.all(got["sad"][:, :, 0] < 0xFFFFFFFF)


def _sad(src, gplane, x0, y0, w, h, dx, dy):
    """SAD of the w x h block at (x0, y0) against the G plane (48 samples of border) at (dx, dy)."""
    blk = gplane[y0 + dy + 48:y0 + dy + 48 + h, x0 + dx + 48:x0 + dx + 48 + w].astype(np.int32)
    return int(np.abs(src[y0:y0 + h, x0:x0 + w].astype(np.int32) - blk).sum())


def test_cost_volume_features_match_numpy():
    """The CNN's input features, restated in numpy: the grid centres from reference 0's co-located
    vector (scaled by the reference's distance, to a multiple of 4 samples, clamped inside the
    macroblock's vectors); channels 0..2 log2(1 + SAD16) / 16 at the 81 cells 4 samples apart (the SAD
    recomputed from the G plane and recovered from the feature); channels 3..5 the cells of the
    co-located, left and above vectors; channel 6 one, channel 7 zero."""
    w, h, nref = 208, 144, 3
    mb_w, mb_h = 13, 9
    pics = frames(w, h, nref + 1, 31, max_motion=12)
    rng = np.random.default_rng(5)
    n = mb_w * mb_h
    mbs = np.zeros(n, dtype=abi.MBINFO_DTYPE)
    mbs["kind"] = np.where(rng.random(n) < 0.7, 5, 1)
    mbs["mv"][:, 5, 0] = rng.integers(-90, 91, n)
    mbs["mv"][:, 5, 1] = rng.integers(-90, 91, n)
    with Analyser(w, h, nrefs=3, search_mode=abi.SEARCH_LEARNED, search_range=32) as an:
        an.load_proposal_weights(proposal.init_params(15))
        for r in range(nref):
            an.set_reference(r, pics[r], mbs if r == 0 else None)
        an.upload(pics[nref])
        feat, ctr = an.proposal_features(list(range(nref)))
        G = [an.read_halfpel(r)[0] for r in range(nref)]
    src = pics[nref].y
    ctr = ctr.reshape(n, 3, 2)
    for mb in range(n):
        mbx, mby = mb % mb_w, mb // mb_w
        x0, y0 = mbx * 16, mby * 16
        lox, hix, loy, hiy = -(x0 + 24), w - x0 + 8, -(y0 + 24), h - y0 + 8
        recs = [(mbx, mby), (mbx - 1, mby), (mbx, mby - 1)]
        tmv = []
        for nx, ny in recs:
            m = mbs[ny * mb_w + nx] if nx >= 0 and ny >= 0 else None
            tmv.append((int(m["mv"][5][0]), int(m["mv"][5][1])) if m is not None and m["kind"] == 5 else None)
        c0 = tmv[0] or (0, 0)
        for r in range(nref):
            px = int(np.floor(((r + 1) * c0[0] + 8) / 16)) * 4
            py = int(np.floor(((r + 1) * c0[1] + 8) / 16)) * 4
            cx, cy = min(max(px, lox + 16), hix - 16), min(max(py, loy + 16), hiy - 16)
            assert (int(ctr[mb, r, 0]), int(ctr[mb, r, 1])) == (cx, cy), (mb, r)
            for c in range(81):
                gx, gy = c % 9, c // 9
                s = _sad(src, G[r], x0, y0, 16, 16, cx + 4 * (gx - 4), cy + 4 * (gy - 4))
                f = float(feat[mb, c, r])
                assert abs(f - np.log2(1.0 + s) / 16) <= 2e-6, (mb, r, c)
                assert abs(2.0 ** (16 * f) - 1 - s) <= 1e-4 * s + 0.05, (mb, r, c, s)
        cx0, cy0 = int(ctr[mb, 0, 0]), int(ctr[mb, 0, 1])
        for c in range(81):
            gx, gy = c % 9, c // 9
            for k in range(3):
                hit = tmv[k] is not None and (int(np.floor((tmv[k][0] - 4 * cx0 + 8) / 16)) + 4,
                                              int(np.floor((tmv[k][1] - 4 * cy0 + 8) / 16)) + 4) == (gx, gy)
                assert feat[mb, c, 3 + k] == (1.0 if hit else 0.0), (mb, c, k)
        assert np.all(feat[mb, :, 6] == 1.0) and np.all(feat[mb, :, 7] == 0.0)


def test_learned_refinement_is_locally_optimal():
    """Every position of a selected cell's 5x5 block is evaluated, so each returned partition winner's
    SAD is at most the SAD at every position that lies in all the blocks (of grid cells around the
    reference's centre) that contain the returned vector; and it is exact at the returned vector."""
    params = proposal.init_params(11)
    w, h, nref = 128, 96, 3
    pics = frames(w, h, nref + 1, 17, max_motion=8)
    with Analyser(w, h, nrefs=3, search_mode=abi.SEARCH_LEARNED, search_range=32, fallback_ratio=1e9) as an:
        an.load_proposal_weights(params)
        for r in range(nref):
            an.set_reference(r, pics[r])
        an.upload(pics[nref])
        an.motion_search(list(range(nref)))
        got = an.read_motion(nref)
        _, ctr = an.proposal_features(list(range(nref)))
        G = [an.read_halfpel(r)[0] for r in range(nref)]
    src = pics[nref].y
    mb_w = w // 16
    ctr = ctr.reshape(-1, 3, 2)
    checked = 0
    for r in range(nref):
        for mb in range(0, (w // 16) * (h // 16), 3):
            mbx, mby = mb % mb_w, mb // mb_w
            cx, cy = int(ctr[mb, r, 0]), int(ctr[mb, r, 1])
            lox, hix, loy, hiy = -(mbx * 16 + 24), w - mbx * 16 + 8, -(mby * 16 + 24), h - mby * 16 + 8
            for p, (x, y, pw_, ph_) in enumerate(abi.PARTITIONS):
                dx, dy = (int(v) // 4 for v in got[r, mb]["mv"][p])
                x0, y0 = mbx * 16 + x * 4, mby * 16 + y * 4
                sad = int(got[r, mb]["sad"][p])
                assert sad == _sad(src, G[r], x0, y0, pw_ * 4, ph_ * 4, dx, dy), (r, mb, p)
                # grid cells whose 5x5 block holds (dx, dy)
                cells = [(gx, gy) for gx in range(9) for gy in range(9)
                         if abs(dx - (cx + 4 * (gx - 4))) <= 2 and abs(dy - (cy + 4 * (gy - 4))) <= 2]
                assert cells, (r, mb, p)
                xs = set.intersection(*[set(range(cx + 4 * (gx - 4) - 2, cx + 4 * (gx - 4) + 3)) for gx, _ in cells])
                ys = set.intersection(*[set(range(cy + 4 * (gy - 4) - 2, cy + 4 * (gy - 4) + 3)) for _, gy in cells])
                for ex in xs:
                    for ey in ys:
                        if lox <= ex <= hix and loy <= ey <= hiy:
                            assert sad <= _sad(src, G[r], x0, y0, pw_ * 4, ph_ * 4, ex, ey), (r, mb, p, ex, ey)
                            checked += 1
    assert checked > 1000
