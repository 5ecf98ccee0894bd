"""The allocation-free loop filter (ia_deblock_picture_noalloc) against ia_promote_reference's.

Two calls in a row through the C ABI on device planes the caller owns, the second on a smaller
picture (the static row counters keep the last call's contents and must start at zero again): each
must equal the filtered reference the existing entry point promotes, byte for byte.
"""
import ctypes as C

import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Analyser, Picture, library

pytestmark = pytest.mark.gpu


def filtered_by_promotion(w, h, seed):
    """An IDR and a P picture analysed; the P picture's reconstruction and records, and its reference
    as ia_promote_reference filters it."""
    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, 2, seed, max_motion=6))]
    with Analyser(w, h, nrefs=1, slices=2) as an:
        idr = an.analyse(pics[0], 30, abi.PIC_IDR)
        an.set_reference(0, idr.recon, idr.mbs)
        res = an.analyse(pics[1], 34, abi.PIC_P, [0])
        an.promote_reference(0, deblock=True)
        ref, _ = an.read_reference(0)
    return res.recon, res.mbs, ref


def noalloc(rec, mbs, mb_w, mb_h):
    import torch

    y = torch.from_numpy(rec.y.copy()).cuda()
    cb = torch.from_numpy(rec.cb.copy()).cuda()
    cr = torch.from_numpy(rec.cr.copy()).cuda()
    m = torch.from_numpy(np.ascontiguousarray(mbs).view(np.uint8).copy()).cuda()
    stalled = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = library().ia_deblock_picture_noalloc(y.data_ptr(), cb.data_ptr(), cr.data_ptr(), m.data_ptr(), mb_w, mb_h,
                                              stalled.data_ptr(), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == abi.IA_OK
    torch.cuda.synchronize()
    return y.cpu().numpy(), cb.cpu().numpy(), cr.cpu().numpy(), int(stalled.item())


@pytest.mark.parametrize("first,second", [((320, 192), (96, 64))])
def test_noalloc_loop_filter_matches_promotion(first, second):
    for (w, h), seed in ((first, 5), (second, 6)):
        rec, mbs, want = filtered_by_promotion(w, h, seed)
        y, cb, cr, stalled = noalloc(rec, mbs, (w + 15) // 16, (h + 15) // 16)
        assert stalled == 0
        assert np.array_equal(y, want.y) and np.array_equal(cb, want.cb) and np.array_equal(cr, want.cr), (w, h)


def test_noalloc_rejects_pictures_beyond_its_scratch():
    lib = library()
    assert lib.ia_deblock_picture_noalloc(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 4, 513, None,
                                          None) == abi.IA_EINVAL
