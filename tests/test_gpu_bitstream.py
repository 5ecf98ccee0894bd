"""The GPU analysis written as a bitstream and decoded: the decoder rebuilds the encoder's references.

The reference's test of its encoder is that its decoder reproduces the encoder's reconstruction
exactly (crates/ironav-codec/src/h264/enc/mod.rs:812-825). Here each picture of a chain is analysed on
the GPU, its macroblock codes are written by the host back end (libironav_h264w), and the picture is
loop-filtered and promoted to a reference on the device (ia_promote_reference) as the encoder keeps
it. The stream is then decoded by the oracle's decoder (pinned by the reference's decoder fixtures):
every decoded picture must equal the promoted reference the GPU holds, byte for byte, and its
macroblock records (kinds, vectors, references, coded-block patterns) the analysis' own.
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen, h264w, proposal
from ironav_amd.h264a import Analyser, Picture

pytestmark = pytest.mark.gpu


def frames(spec):
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


def chain(spec, *, nrefs=3, slices=8, search_mode=abi.SEARCH_FULL, search_range=32, qp=28, fast=False,
          params=None, transform8=True, kinds=None):
    from oracle import pydecoder

    pics = frames(spec)
    w, h = spec.width, spec.height
    stream = bytearray()
    refs_dev, mbs_dev = [], []
    with Analyser(w, h, nrefs=nrefs, slices=slices, search_mode=search_mode, search_range=search_range,
                  fast=fast, transform8=transform8) as an, \
            h264w.Stream(w, h, nrefs=nrefs, slices=slices, transform8=transform8) as wr:
        if params is not None:
            an.load_proposal_weights(params)
        order = []  # slots, newest first
        for t, pic in enumerate(pics):
            kind = kinds[t] if kinds else (abi.PIC_IDR if t == 0 else abi.PIC_P)
            if kind == abi.PIC_IDR:
                order = []
            q = qp - 3 if kind == abi.PIC_IDR else qp
            refs = order[:nrefs] if kind == abi.PIC_P else []
            res = an.analyse(pic, q, kind, refs)
            stream += wr.write(kind, q, res.codes, len(refs))
            slot = order.pop() if len(order) == nrefs else len(order)
            an.promote_reference(slot, deblock=True)
            order.insert(0, slot)
            ref, _ = an.read_reference(slot)
            refs_dev.append(ref)
            mbs_dev.append(res.mbs)
    got = pydecoder.decode_annexb(bytes(stream))
    assert len(got) == len(pics)
    for t, (f, ref, mbs) in enumerate(zip(got, refs_dev, mbs_dev)):
        for k, name in enumerate(("y", "cb", "cr")):
            assert np.array_equal(f.full[k], getattr(ref, name)), f"picture {t}: decoded {name} differs from the reference"
        for fld in ("kind", "cbp", "skip", "t8x8", "mv", "ref_idx"):
            assert np.array_equal(f.mbs[fld], mbs[fld]), f"picture {t}: decoded {fld} differs"
    return len(stream)


def test_1080p_workload_chain_decodes_to_the_references():
    """1920x1080, IDR + 4 P pictures, 3 references, +-32 exhaustive search, 8 slices (the encoder's
    rule on 8 or more host threads), CABAC + 8x8 transform."""
    n = chain(clipgen.ClipSpec(1920, 1080, 5, 1080, max_motion=12, noise_sigma=2.0))
    print(f"1080p chain: {n} bytes for 5 pictures")


def test_learned_proposal_chain_decodes():
    """The learned proposal's winners as the analysis' extra candidates (trained weights, fallback
    1.1x), 640x360, with a scene cut coded as an IDR picture."""
    spec = clipgen.ClipSpec(640, 360, 5, 77, max_motion=24, cuts=(3,))
    pics = frames(spec)
    with h264w.Gop(640, 360) as g:
        kinds = [g.next(p.y) for p in pics]
    assert kinds == [abi.PIC_IDR, abi.PIC_P, abi.PIC_P, abi.PIC_IDR, abi.PIC_P]
    chain(spec, slices=4, search_mode=abi.SEARCH_LEARNED, params=proposal.load_weights(), kinds=kinds)


def test_realtime_and_main_profile_chains_decode():
    chain(clipgen.ClipSpec(320, 192, 4, 61, max_motion=8), nrefs=1, slices=6, search_mode=abi.SEARCH_DIAMOND,
          fast=True)
    chain(clipgen.ClipSpec(208, 120, 4, 62, max_motion=8), nrefs=2, slices=2, search_mode=abi.SEARCH_DIAMOND,
          transform8=False, qp=34)


def test_2160p_large_motion_chain_decodes():
    """3840x2160, objects +-64, noise 4, +-64 exhaustive search, 2 references, 34 slices."""
    chain(clipgen.ClipSpec(3840, 2160, 3, 8002, max_motion=64, noise_sigma=4.0), nrefs=2, slices=34,
          search_range=64)
