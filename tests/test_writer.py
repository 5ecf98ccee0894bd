"""The bitstream back end (ironav_amd/csrc/writer.c) on the CPU.

Pinned by the reference's own tests, restated with their inputs and expected outputs:
- the exact bytes of a NAL unit with emulation prevention (crates/ironav-codec/src/h264/enc/bitw.rs:137-142);
- the SPS level for five sizes and rates (enc/mod.rs:1057-1069);
- the CABAC coder against the decoder's arithmetic decoder: 20,000 seeded bins (decisions, bypass bins,
  terminate bins) read back, and the flush ending the data (enc/cabac.rs:252-290);
- the system invariant of enc/mod.rs:812-825: the decoder rebuilds the encoder's reconstruction. Here
  the oracle's analysis of IPPP chains (its references loop-filtered, as the encoder keeps them) is
  written by the back end and decoded by the oracle's decoder (itself pinned by the reference's
  decoder fixtures): every decoded picture equals the loop-filtered reconstruction byte for byte, and
  the decoded macroblock records (kinds, vectors, references, coded-block patterns) equal the
  analysis'. The written size stays within 2% of the rate the analysis counted with its CABAC model
  (ia_mbstats.bits) plus headers.
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen, h264w
from ironav_amd.h264a import Picture


def test_nal_emulation_prevention_bytes():
    out = h264w.put_nal(3, 5, bytes([0, 0, 0, 0, 0, 1, 0, 0, 3, 7]))
    assert list(out) == [0, 0, 0, 1, 0x65, 0, 0, 3, 0, 0, 3, 0, 1, 0, 0, 3, 3, 7]


def test_sps_level_follows_size_and_rate():
    assert h264w.level(1280, 720, 30, 1) == 31
    assert h264w.level(1920, 1080, 30, 1) == 40
    assert h264w.level(1920, 1080, 60, 1) == 42
    assert h264w.level(2560, 1440, 30, 1) == 50
    assert h264w.level(3840, 2160, 30, 1) == 51


def test_cabac_bins_read_back_through_the_decoder(oracle):
    from oracle import pydecoder

    seed = 4321

    def rnd(m):
        nonlocal seed
        seed = (seed * 1103515245 + 12345) & 0xFFFFFFFF
        return (seed >> 16) % m

    kinds, ctx, vals = [], [], []
    for _ in range(20000):
        k = [0, 0, 0, 1, 2][rnd(5)]
        c = rnd(8) * 37
        v = rnd(10) < (2 if c % 2 == 0 else 8)
        kinds.append(k)
        ctx.append(c)
        vals.append(int(v) if k != 2 else 0)
    for intra, qp in ((False, 30), (True, 12)):
        data = h264w.cabac_encode_bins(kinds, ctx, vals, intra, qp)
        got, fin, nread = pydecoder.cabac_decode_bins(data, kinds, ctx, intra, qp)
        assert list(got) == vals
        assert fin == 1
        assert nread == len(data), "the flush ends the data"


def frames(w, h, n, seed, **kw):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, n, seed, **kw))]


def chain_roundtrip(oracle, w, h, nframes, qp, *, nrefs=3, slices=1, transform8=True, sub8x8=False,
                    search_mode=abi.SEARCH_DIAMOND, search_range=16, fast=False, seed=1, **clip):
    from oracle import pydecoder

    pics = frames(w, h, nframes, seed, **clip)
    cfg = abi.Config(w, h, 8, 1, int(transform8), int(sub8x8), int(fast), nrefs, slices, search_mode, search_range, 1.1)
    mb_w, mb_h = (w + 15) // 16, (h + 15) // 16
    stream = bytearray()
    want = []
    refs = []
    model_bits = []
    with h264w.Stream(w, h, nrefs=nrefs, slices=slices, transform8=transform8) as wr:
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            q = qp - 3 if t == 0 else qp
            nref = min(len(refs), nrefs) if t else 0
            fs = None
            if kind == abi.PIC_P and search_mode == abi.SEARCH_FULL:
                fs = oracle.fullsearch(pic.y, [r[0].y for r in refs[:nref]], mb_w, mb_h, search_range)
            codes, stats, rec, mbs = oracle.analyse_picture(cfg, pic, q, kind, refs[:nref], fs)
            au = wr.write(kind, q, codes, nref)
            stream += au
            model_bits.append((len(au) * 8, int(stats["bits"].sum()) / 256.0))
            filt = oracle.deblock_picture(rec[0], rec[1], rec[2], mbs, mb_w, mb_h)
            want.append((filt, codes, mbs))
            refs.insert(0, (Picture(*filt), mbs))
    got = pydecoder.decode_annexb(bytes(stream))
    assert len(got) == nframes
    for t, (f, (filt, codes, mbs)) in enumerate(zip(got, want)):
        for k in range(3):
            assert np.array_equal(f.full[k], filt[k]), f"picture {t}: decoded plane {k} differs from the reconstruction"
        # the cropped output is the reconstruction's visible part
        assert np.array_equal(f.y, filt[0][:h, :w])
        for fld in ("kind", "cbp", "skip", "t8x8", "mv", "ref_idx"):
            assert np.array_equal(f.mbs[fld], mbs[fld]), f"picture {t}: decoded {fld} differs"
    for t, (actual, model) in enumerate(model_bits):
        # headers: SPS/PPS, and per slice its start code, NAL header, slice header, alignment and flush
        assert abs(actual - model) <= 0.02 * model + 128 * slices + (400 if t == 0 else 0), (t, actual, model)
    return got


@pytest.mark.parametrize("qp,slices,nrefs", [(22, 1, 3), (28, 2, 2), (36, 3, 1)])
def test_oracle_chain_decodes_to_the_reconstruction(oracle, qp, slices, nrefs):
    chain_roundtrip(oracle, 96, 64, 4, qp, nrefs=nrefs, slices=slices, seed=qp)


def test_full_search_sub8x8_and_main_profile(oracle):
    chain_roundtrip(oracle, 64, 48, 3, 18, nrefs=2, slices=1, sub8x8=True, search_mode=abi.SEARCH_FULL,
                    search_range=12, seed=4, max_motion=8)
    chain_roundtrip(oracle, 80, 48, 3, 30, nrefs=1, slices=1, transform8=False, seed=8)


def test_cropped_size_and_realtime_decisions(oracle):
    chain_roundtrip(oracle, 40, 26, 3, 24, nrefs=1, slices=2, seed=3)
    chain_roundtrip(oracle, 96, 64, 3, 30, nrefs=1, slices=2, fast=True, seed=60, max_motion=2)


def test_extreme_qps(oracle):
    chain_roundtrip(oracle, 48, 32, 2, 51, nrefs=1, seed=12)
    chain_roundtrip(oracle, 48, 32, 2, 3, nrefs=1, seed=13)


def test_cavlc_is_refused():
    with pytest.raises(NotImplementedError):
        h264w.Stream(64, 48, cabac=False)


def b_chain_roundtrip(oracle, w, h, nframes, qp, *, nrefs=3, slices=1, transform8=True, idr=None, seed=1, **clip):
    """An IBBP chain as the reference's encoder codes it with two B pictures (enc/mod.rs:700-731):
    pictures in coding order (h264w.coding_order), P pictures against the references kept (newest
    first), B pictures against the nearest reference shown before and after them, QP offsets -3 / 0 /
    +5 (enc/mod.rs:36); each analysed by the oracle, written by the back end, loop-filtered; the stream
    decoded by the oracle's decoder must give every picture (B pictures too) and its records."""
    from oracle import pydecoder

    pics = frames(w, h, nframes, seed, **clip)
    cfg = abi.Config(w, h, 8, 1, int(transform8), 0, 0, nrefs, slices, abi.SEARCH_DIAMOND, 16, 1.1)
    mb_w, mb_h = (w + 15) // 16, (h + 15) // 16
    idr = idr if idr is not None else [d == 0 for d in range(nframes)]
    order = h264w.coding_order(idr, 2)
    ref_count = max(nrefs, 2)
    refs, want, stream, model_bits = [], {}, bytearray(), []
    with h264w.Stream(w, h, nrefs=nrefs, slices=slices, transform8=transform8, bframes=2) as wr:
        for d, kind in order:
            pic = pics[d]
            if kind == abi.PIC_IDR:
                refs.clear()
                codes, stats, rec, mbs = oracle.analyse_picture(cfg, pic, qp - 3, kind)
                au = wr.write_at(kind, qp - 3, codes, d)
            elif kind == abi.PIC_P:
                plist = refs[::-1][:nrefs]
                codes, stats, rec, mbs = oracle.analyse_picture(cfg, pic, qp, kind, [(r[1], r[2]) for r in plist])
                au = wr.write_at(kind, qp, codes, d, len(plist))
            else:
                l0 = max((r for r in refs if r[0] < d), key=lambda r: r[0])
                l1 = min((r for r in refs if r[0] > d), key=lambda r: r[0])
                tb, td = 2 * (d - l0[0]), 2 * (l1[0] - l0[0])
                codes, stats, rec, mbs = oracle.analyse_picture_b(cfg, pic, min(qp + 5, 51), (l0[1], l0[2]),
                                                                  (l1[1], l1[2]), tb, td)
                au = wr.write_at(kind, min(qp + 5, 51), codes, d)
            stream += au
            model_bits.append((d, kind, len(au) * 8, int(stats["bits"].sum()) / 256.0))
            filt = oracle.deblock_picture(rec[0], rec[1], rec[2], mbs, mb_w, mb_h)
            want[d] = (filt, codes, mbs, kind)
            if kind != abi.PIC_B:
                if len(refs) == ref_count:
                    refs.pop(0)
                refs.append((d, Picture(*filt), mbs))
    got = pydecoder.decode_annexb(bytes(stream))
    assert len(got) == nframes
    for d, f in enumerate(got):
        filt, codes, mbs, kind = want[d]
        for k in range(3):
            assert np.array_equal(f.full[k], filt[k]), f"picture {d} (kind {kind}): decoded plane {k} differs"
        flds = ("kind", "cbp", "skip", "t8x8", "mv", "ref_idx") + (("mv1", "ref_idx1", "bdirect", "direct8") if kind == abi.PIC_B else ())
        for fld in flds:
            assert np.array_equal(f.mbs[fld], mbs[fld]), f"picture {d} (kind {kind}): decoded {fld} differs"
    for d, kind, actual, model in model_bits:
        assert abs(actual - model) <= 0.02 * model + 128 * slices + (400 if kind == abi.PIC_IDR else 0), (d, kind, actual, model)
    return got, want


def test_coding_order_with_b_pictures():
    """enc/mod.rs:1110-1159 with two B pictures: IDR pictures at 0, 3 (forced) and 8 (keyint 5); every
    picture is coded once, each B picture after the P or IDR picture shown next."""
    idr = [d in (0, 3, 8) for d in range(12)]
    order = h264w.coding_order(idr, 2)
    assert sorted(d for d, _ in order) == list(range(12))
    assert [d for d, k in order if k == abi.PIC_IDR] == [0, 3, 8]
    assert order == [(0, 0), (2, 1), (1, 2), (3, 0), (6, 1), (4, 2), (5, 2), (7, 1), (8, 0), (11, 1), (9, 2), (10, 2)]
    assert h264w.coding_order([True] + [False] * 6, 0) == [(0, 0)] + [(d, 1) for d in range(1, 7)]


@pytest.mark.parametrize("qp,slices,nrefs", [(24, 1, 3), (30, 2, 1), (36, 1, 2)])
def test_oracle_b_chain_decodes_to_the_reconstruction(oracle, qp, slices, nrefs):
    got, want = b_chain_roundtrip(oracle, 96, 64, 8, qp, nrefs=nrefs, slices=slices, seed=qp + 100)
    # the B pictures used each coding form
    codes = np.concatenate([want[d][1] for d in want if want[d][3] == abi.PIC_B])
    inter = codes[codes["kind"] == abi.KIND_INTER]
    assert (inter["skip"] == 1).any()


def test_b_chain_main_profile_and_cropped(oracle):
    b_chain_roundtrip(oracle, 40, 26, 6, 26, nrefs=1, slices=2, transform8=False, seed=7)
    b_chain_roundtrip(oracle, 64, 48, 7, 20, nrefs=2, slices=1, seed=9, max_motion=6)
