"""GPU parity of the loop filter and of the on-device reference promotion, against the oracle.

`ia_promote_reference` takes the last analysed picture to a reference slot on the device, loop-
filtered as the encoder filters each picture before keeping it (enc/mod.rs:653-667). Checked byte for
byte: the slot's luma, chroma and records against `or_deblock_picture` of the analysis's own
reconstruction and records, and the slot's half-sample planes against the oracle's planes of that
filtered luma. The chains run the encoder's whole picture loop on the GPU (analyse, filter, promote,
search the next picture against the filtered references) beside the same loop in the oracle, and
compare every picture's analysis too.
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Analyser, Picture

pytestmark = pytest.mark.gpu


def frames(w, h, n, seed, **kw):
    spec = clipgen.ClipSpec(w, h, n, seed, **kw)
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


def chain(oracle, w, h, nframes, qp, *, nrefs=2, slices=1, search_mode=abi.SEARCH_DIAMOND, search_range=16,
          seed=1, deblock=True, **clip):
    pics = frames(w, h, nframes, seed, **clip)
    filtered = 0
    with Analyser(w, h, nrefs=nrefs, slices=slices, search_mode=search_mode, search_range=search_range) as an:
        order = []  # slots, newest picture first
        refs = []   # the oracle's references (Picture, mbinfo), newest first
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            q = qp - 3 if kind == abi.PIC_IDR else qp
            nref = min(len(order), nrefs)
            slots = order[:nref]
            fs = None
            if kind == abi.PIC_P and search_mode == abi.SEARCH_FULL:
                an.upload(pic)
                an.motion_search(slots)
                fs = an.read_motion(nref)
            res = an.analyse(pic, q, kind, slots)
            codes, stats, (ry, rcb, rcr), mbs = oracle.analyse_picture(an.cfg, pic, q, kind, refs[:nref], fs)
            assert np.array_equal(res.codes, codes) and np.array_equal(res.mbs, mbs), f"picture {t}: analysis"
            assert np.array_equal(res.recon.y, ry) and np.array_equal(res.recon.cb, rcb), f"picture {t}: recon"
            # promote into the oldest slot (or a free one)
            slot = order.pop() if len(order) == nrefs else len(order)
            an.promote_reference(slot, deblock=deblock)
            order.insert(0, slot)
            got, gmbs = an.read_reference(slot)
            if deblock:
                want = oracle.deblock_picture(ry, rcb, rcr, mbs, an.mb_w, an.mb_h)
            else:
                want = (ry, rcb, rcr)
            for k, name in enumerate(("y", "cb", "cr")):
                g, wv = (got.y, got.cb, got.cr)[k], want[k]
                if not np.array_equal(g, wv):
                    bad = np.argwhere(g != wv)
                    raise AssertionError(f"picture {t} qp {q}: {name} differs at {len(bad)} samples, first "
                                         f"{tuple(bad[0])}: {g[tuple(bad[0])]} vs {wv[tuple(bad[0])]}")
            assert np.array_equal(gmbs, mbs)
            filtered += int(not np.array_equal(want[0], ry))
            hp = an.read_halfpel(slot)
            whp = oracle.halfpel_build(want[0].astype(np.uint16), 8)
            for k in range(4):
                assert np.array_equal(hp[k], whp[k].astype(np.uint8)), f"picture {t}: half-sample plane {'GBHJ'[k]}"
            refs.insert(0, (Picture(*want), mbs))
    return filtered


@pytest.mark.parametrize("qp", [24, 32, 40])
def test_chain_with_loop_filter(oracle, qp):
    assert chain(oracle, 128, 96, 4, qp, nrefs=2, slices=2, seed=qp) > 0


def test_chain_single_and_odd_sizes(oracle):
    chain(oracle, 16, 16, 3, 36, nrefs=1, seed=3)
    chain(oracle, 40, 26, 3, 30, nrefs=2, slices=2, seed=4)
    chain(oracle, 208, 48, 3, 44, nrefs=1, seed=5, max_motion=12, noise_sigma=3.0)


def test_promotion_without_filter_is_the_reconstruction(oracle):
    assert chain(oracle, 96, 64, 3, 34, nrefs=2, seed=7, deblock=False) == 0


def test_1080p_chain_full_search(oracle):
    """1920x1080, IDR + 2 P pictures, 3 references, 8 slices, exhaustive search +-32: the encoder's
    picture loop with every reference filtered and promoted on the GPU."""
    assert chain(oracle, 1920, 1080, 3, 30, nrefs=3, slices=8, search_mode=abi.SEARCH_FULL, search_range=32,
                 seed=1080) > 0


def test_repeated_promotion_is_stable(oracle):
    """The row counters return to 0 after every launch: promoting the same picture again gives the
    same slot contents."""
    pics = frames(64, 48, 1, 11)
    with Analyser(64, 48, nrefs=2) as an:
        an.analyse(pics[0], 30, abi.PIC_IDR)
        an.promote_reference(0)
        a = an.read_reference(0)
        for _ in range(3):
            an.promote_reference(1)
            b = an.read_reference(1)
            assert all(np.array_equal(u, v) for u, v in zip((a[0].y, a[0].cb, a[0].cr), (b[0].y, b[0].cb, b[0].cr)))
