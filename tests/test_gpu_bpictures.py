"""B pictures on the GPU (ia_analyse_b) against the oracle, and written, decoded and compared.

Chains are coded as the reference's encoder codes them with two B pictures (enc/mod.rs:700-731,
h264w.coding_order): IDR and P pictures analysed and promoted to loop-filtered references on the
device, each B picture analysed against the nearest reference shown before it (list 0) and after it
(list 1), whose records give the co-located motion of the spatial direct prediction. Every B
picture's codes, costs, reconstruction and records must equal the oracle's (oracle.analyse_picture_b,
enc/frame.rs:1518-1780) on the same references, bit for bit; P pictures likewise. The whole chain is
then written by the host back end and decoded by the oracle's decoder (pinned by the reference's B
fixtures): every picture, B pictures included, decodes to the loop-filtered reconstruction.
"""
import numpy as np
import pytest

from ironav_amd import abi, clipgen, h264w
from ironav_amd.h264a import Analyser, Picture

from test_gpu_analyse import compare

pytestmark = pytest.mark.gpu


def b_chain(oracle, spec, qp, *, nrefs=2, slices=1, transform8=True, threads=8, check_every_p=True):
    from oracle import pydecoder

    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]
    w, h = spec.width, spec.height
    ref_count = max(nrefs, 2)
    order = h264w.coding_order([d == 0 for d in range(len(pics))], 2)
    refs = []          # (display, slot) in coding order
    free = list(range(ref_count))
    want_pics, nb = {}, 0
    stream = bytearray()
    with Analyser(w, h, nrefs=ref_count, slices=slices, transform8=transform8) as an, \
            h264w.Stream(w, h, nrefs=nrefs, slices=slices, transform8=transform8, bframes=2) as wr:
        mb_w, mb_h = an.mb_w, an.mb_h
        for d, kind in order:
            pic = pics[d]
            print(f"picture {d} kind {kind}", flush=True)
            if kind == abi.PIC_IDR:
                free = list(range(ref_count))
                refs = []
                q = qp - 3
                res = an.analyse(pic, q, kind)
                want = oracle.analyse_picture(an.cfg, pic, q, kind)
                compare(f"IDR {d}", (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs), want)
                stream += wr.write_at(kind, q, res.codes, d)
            elif kind == abi.PIC_P:
                q = qp
                plist = refs[::-1][:nrefs]
                slots = [s for _, s in plist]
                res = an.analyse(pic, q, kind, slots)
                if check_every_p:
                    rl = [an.read_reference(s) for s in slots]
                    want = oracle.analyse_picture(an.cfg, pic, q, kind, rl, threads=threads)
                    compare(f"P {d}", (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs), want)
                stream += wr.write_at(kind, q, res.codes, d, len(slots))
            else:
                q = min(qp + 5, 51)
                l0 = max((r for r in refs if r[0] < d), key=lambda r: r[0])
                l1 = min((r for r in refs if r[0] > d), key=lambda r: r[0])
                tb, td = 2 * (d - l0[0]), 2 * (l1[0] - l0[0])
                res = an.analyse_b(pic, q, l0[1], l1[1], tb, td)
                r0, r1 = an.read_reference(l0[1]), an.read_reference(l1[1])
                want = oracle.analyse_picture_b(an.cfg, pic, q, r0, r1, tb, td, threads=threads)
                compare(f"B {d}", (res.codes, res.stats, (res.recon.y, res.recon.cb, res.recon.cr), res.mbs), want)
                stream += wr.write_at(kind, q, res.codes, d)
                want_pics[d] = oracle.deblock_picture(res.recon.y, res.recon.cb, res.recon.cr, res.mbs, mb_w, mb_h)
                nb += 1
                with pytest.raises(Exception):
                    an.promote_reference(l0[1])  # a B picture is not a reference
                continue
            if len(refs) == ref_count:
                _, slot = refs.pop(0)
            else:
                slot = free.pop(0)
            an.promote_reference(slot, deblock=True)
            refs.append((d, slot))
            ref, _ = an.read_reference(slot)
            want_pics[d] = (ref.y, ref.cb, ref.cr)
    assert nb > 0
    got = pydecoder.decode_annexb(bytes(stream))
    assert len(got) == len(pics)
    for d, f in enumerate(got):
        for k, name in enumerate(("y", "cb", "cr")):
            assert np.array_equal(f.full[k], want_pics[d][k]), f"picture {d}: decoded {name} differs"
    return len(stream)


@pytest.mark.parametrize("qp,slices,nrefs", [(24, 1, 2), (30, 2, 1), (36, 3, 3)])
def test_b_chain_against_the_oracle(oracle, qp, slices, nrefs):
    b_chain(oracle, clipgen.ClipSpec(96, 64, 7, 200 + qp, max_motion=8), qp, nrefs=nrefs, slices=slices)


def test_b_chain_main_profile_odd_size(oracle):
    b_chain(oracle, clipgen.ClipSpec(208, 120, 5, 207, max_motion=10), 28, nrefs=2, slices=2, transform8=False)


def test_b_chain_1080p(oracle):
    """1920x1080, I P B B P B B, 3 references for P pictures, 8 slices, CABAC + 8x8 transform."""
    n = b_chain(oracle, clipgen.ClipSpec(1920, 1080, 7, 1081, max_motion=12, noise_sigma=2.0), 28, nrefs=3,
                slices=8, check_every_p=False)
    print(f"1080p IBBP chain: {n} bytes for 7 pictures")
