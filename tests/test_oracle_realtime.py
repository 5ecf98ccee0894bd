"""The oracle's realtime decisions (enc/frame.rs:731-765, 861-863, 973, 990-992, 1006), pinned by what
they must do to a picture, each on inputs where the default decisions do otherwise:

* a P macroblock within a twentieth of a quantiser step of its skip prediction, in luma and chroma,
  is P_Skip with the prediction as its reconstruction, even where a transform would code levels;
* P macroblocks are 16x16 only, with vectors on half samples (no quarter-sample pass);
* in I pictures coded with CAVLC, a macroblock whose Intra16x16 cost is below 256 * 0.8 * 2^(QP/6)
  is never Intra4x4.

The shorter diamond ({8, 4} quarter samples in place of {16, 8, 4}) and the stop at an integer
vector are pinned only through the GPU parity tests: no property of the output fixes them.
"""
import numpy as np

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Picture


def _cfg(w, h, *, cabac, fast, nrefs=1, slices=1):
    return abi.Config(w, h, 8, int(cabac), int(cabac), 0, int(fast), nrefs, slices, abi.SEARCH_DIAMOND, 16, 1.1)


def _clip(w, h, n, seed, **kw):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, n, seed, **kw))]


def _mb(plane, mbx, mby, n):
    return plane[mby * n:(mby + 1) * n, mbx * n:(mbx + 1) * n]


def test_near_copies_are_skipped_without_a_transform(oracle):
    w, h, qp = 64, 48, 10
    mb_w, mb_h = w // 16, h // 16
    src0 = _clip(w, h, 1, 7)[0]
    _, _, rec, mbs = oracle.analyse_picture(_cfg(w, h, cabac=False, fast=False), src0, qp, abi.PIC_IDR)
    ref = Picture(*rec)
    y, cb, cr = ref.y.astype(np.int32), ref.cb.astype(np.int32), ref.cr.astype(np.int32)
    # luma 4 samples x 3 = SAD 12 per macroblock, under 256 * 0.05 * 2^(10/6) = 40.6
    for mby in range(mb_h):
        for mbx in range(mb_w):
            for k in range(4):
                yy, xx = mby * 16 + 3 + 4 * k, mbx * 16 + 2 + 3 * k
                y[yy, xx] = y[yy, xx] + 3 if y[yy, xx] < 250 else y[yy, xx] - 3
    # one macroblock's chroma 4 samples x 6 = SAD 24, over 128 * 0.05 * 2^(10/6) = 20.3
    far_x, far_y = 2, 1
    for k in range(4):
        yy, xx = far_y * 8 + 1 + 2 * k, far_x * 8 + 1 + k
        cb[yy, xx] = cb[yy, xx] + 6 if cb[yy, xx] < 240 else cb[yy, xx] - 6
    src = Picture(np.clip(y, 0, 255), np.clip(cb, 0, 255), cr)
    runs = {}
    for fast in (False, True):
        runs[fast] = oracle.analyse_picture(_cfg(w, h, cabac=False, fast=fast), src, qp, abi.PIC_P, [(ref, mbs)])
    codes, stats, rrec, _ = runs[True]
    far = far_y * mb_w + far_x
    for i in range(mb_w * mb_h):
        if i == far:
            continue
        assert codes["skip"][i] == 1 and codes["kind"][i] == abi.KIND_INTER, i
        assert not codes["mvs"][i].any() and codes["cbp"][i] == 0
        mbx, mby = i % mb_w, i // mb_w
        assert np.array_equal(_mb(rrec[0], mbx, mby, 16), _mb(ref.y, mbx, mby, 16)), i
        assert np.array_equal(_mb(rrec[1], mbx, mby, 8), _mb(ref.cb, mbx, mby, 8)), i
    assert stats["skip_sad"][0] == 12
    # the same picture without the realtime decisions codes the differences somewhere, and the
    # macroblock whose chroma is too far is not skipped either way
    assert (runs[False][0]["skip"] == 0).sum() > 0
    assert codes["skip"][far] == 0


def test_p_pictures_are_16x16_on_half_samples(oracle):
    w, h = 96, 64
    pics = _clip(w, h, 3, 31, max_motion=12, noise_sigma=2.0)
    parts, odd = {}, {}
    for fast in (False, True):
        cfg = _cfg(w, h, cabac=True, fast=fast, nrefs=2)
        refs = []
        part_seen, odd_seen = 0, 0
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            codes, stats, rec, mbs = oracle.analyse_picture(cfg, pic, 26, kind, refs[:2])
            coded = (codes["kind"] == abi.KIND_INTER) & (codes["skip"] == 0)
            part_seen += int((codes["part"][coded] != 0).sum())
            odd_seen += int((codes["mvs"][coded][:, 0, :] & 1).any(axis=1).sum())
            if fast:
                assert (stats["inter_part"][stats["inter_part"] >= 0] == 0).all()
            refs.insert(0, (Picture(*rec), mbs))
        parts[fast], odd[fast] = part_seen, odd_seen
    assert parts[False] > 0 and odd[False] > 0, "the default decisions must use both for the check to bite"
    assert parts[True] == 0 and odd[True] == 0


def test_flat_intra_stays_16x16_with_cavlc(oracle):
    w, h, qp = 96, 64, 24
    pic = _clip(w, h, 1, 17)[0]
    thresh = 256.0 * 0.8 * 2.0 ** (qp / 6.0)
    out = {}
    for fast in (False, True):
        codes, stats, _, _ = oracle.analyse_picture(_cfg(w, h, cabac=False, fast=fast), pic, qp, abi.PIC_IDR)
        out[fast] = (codes, stats)
    codes, stats = out[True]
    flat = stats["i16_cost"] < thresh
    assert flat.any() and (~flat).any()
    assert (codes["kind"][flat] != abi.KIND_I4X4).all()
    # without the rule some of those macroblocks take Intra4x4, so the rule is what keeps them out
    assert (out[False][0]["kind"][flat] == abi.KIND_I4X4).any()
    # and the rule leaves the others free to take it
    assert (codes["kind"][~flat] == abi.KIND_I4X4).any()
