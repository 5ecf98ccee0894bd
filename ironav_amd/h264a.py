"""Host side of the MI355X H.264 analysis stage: the reference's slice-coder interface over the C ABI.

`Analyser` stands where nervosys/IronAV's `code_picture` builds one `FrameCoder` per slice and runs
`code_mb` over every macroblock (crates/ironav-codec/src/h264/enc/mod.rs:603-628,
crates/ironav-codec/src/h264/enc/frame.rs:381-487): it takes the same inputs (the padded source, the
list-0 references with their half-sample planes, the slice QP, the slice bands, the CABAC /
8x8-transform / sub-partition options) and returns what the slice loop hands to the entropy coder
(the `MbCode` of every macroblock, enc/frame.rs:275-310), the reconstruction before deblocking and
the `MbInfo` records (h264/mb.rs:34-68), plus the per-slice counts [intra, inter, skip]
(enc/frame.rs:374,2322).

Errors: invalid arguments raise ValueError (the reference's `H264Encoder::new`/`source` return
`Error::invalid_input`, enc/mod.rs:137-153,483-495); device failures raise RuntimeError. The GPU
library must be built (`python -m ironav_amd.build`); nothing here falls back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libironav_h264a.so")


def batch_workers():
    """Analysis workgroups resident at once on the current device (the size of a batch's task pool)."""
    return int(library().ia_batch_workers())


def library():
    """The loaded gfx950 library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ironav_amd.build`")
        try:
            # one HIP runtime per process: torch ships its own copy under a different file name, so
            # when torch is present it is loaded first and the library binds to that runtime
            import torch  # noqa: F401
        except ImportError:
            pass
        _LIB = abi.bind(C.CDLL(LIB_PATH))
    return _LIB


class AnalysisError(RuntimeError):
    pass


def _check(rc, ctx=None):
    if rc == abi.IA_OK:
        return
    msg = library().ia_last_error(ctx).decode(errors="replace")
    if rc in (abi.IA_EINVAL,):
        raise ValueError(msg)
    raise AnalysisError(f"error {rc}: {msg}")


class Picture:
    """A padded 8-bit 4:2:0 picture: luma mb_w*16 x mb_h*16, chroma half that."""

    def __init__(self, y, cb, cr):
        self.y = np.ascontiguousarray(y, dtype=np.uint8)
        self.cb = np.ascontiguousarray(cb, dtype=np.uint8)
        self.cr = np.ascontiguousarray(cr, dtype=np.uint8)
        if self.y.shape[0] % 16 or self.y.shape[1] % 16:
            raise ValueError("luma must be padded to whole macroblocks")
        if self.cb.shape != (self.y.shape[0] // 2, self.y.shape[1] // 2) or self.cr.shape != self.cb.shape:
            raise ValueError("chroma planes must be half the luma size")


class Analysis:
    """The results of analysing one picture."""

    def __init__(self, codes, stats, recon, mbs, mb_w, mb_h, rows):
        self.codes = codes      # ia_mbcode per macroblock, raster order
        self.stats = stats      # ia_mbstats per macroblock
        self.recon = recon      # Picture: reconstruction before deblocking
        self.mbs = mbs          # ia_mbinfo per macroblock
        self.mb_w, self.mb_h, self.rows = mb_w, mb_h, rows

    def counts(self):
        """[intra, inter, skip] per slice, as each `FrameCoder` counts them."""
        out = []
        for r0, r1 in self.rows:
            c = self.codes[r0 * self.mb_w:r1 * self.mb_w]
            skip = int(np.count_nonzero(c["skip"]))
            inter = int(np.count_nonzero(c["kind"] == abi.KIND_INTER)) - skip
            out.append([len(c) - inter - skip, inter, skip])
        return out


def slice_rows(mb_h, slices):
    """The slice bands of enc/mod.rs:603."""
    n = max(1, min(slices, mb_h))
    return [(i * mb_h // n, (i + 1) * mb_h // n) for i in range(n)]


def realtime(height, threads):
    """The Analyser settings of the encoder's live mode (enc/mod.rs:249-254, `realtime()`): the
    realtime decisions, one reference, and a slice per host thread up to one per two macroblock rows.
    Pass them as keyword arguments: Analyser(w, h, **realtime(h, threads))."""
    mb_h = (height + 15) // 16
    return dict(fast=True, nrefs=1, slices=max(1, min(threads, mb_h // 2)))


class Analyser:
    """One encoder's analysis context on the current HIP device."""

    def __init__(self, width, height, *, nrefs=3, slices=1, cabac=True, transform8=True, sub8x8=False,
                 search_mode=abi.SEARCH_DIAMOND, search_range=32, fallback_ratio=1.1, fast=False):
        self.cfg = abi.Config(width, height, 8, int(cabac), int(transform8), int(sub8x8), int(fast), nrefs, slices,
                              search_mode, search_range, fallback_ratio)
        lib = library()
        h = C.c_void_p()
        _check(lib.ia_create(C.byref(h), C.byref(self.cfg)))
        self._h = h
        self.mb_w, self.mb_h = (width + 15) // 16, (height + 15) // 16
        self.pw, self.ph = self.mb_w * 16, self.mb_h * 16
        self.rows = slice_rows(self.mb_h, slices)

    def close(self):
        if getattr(self, "_h", None):
            library().ia_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, rc):
        _check(rc, self._h)

    def set_stream(self, stream_handle):
        """Enqueue on a caller's HIP stream (an integer handle; 0 for the context's own)."""
        self._ok(library().ia_set_stream(self._h, stream_handle or None))

    def _pic(self, pic):
        if pic.y.shape != (self.ph, self.pw):
            raise ValueError(f"picture is {pic.y.shape}, expected {(self.ph, self.pw)}")
        return pic

    def set_reference(self, slot, pic, mbs=None):
        """Loads a deblocked reference picture (and its macroblock records) into `slot`."""
        pic = self._pic(pic)
        if mbs is not None:
            mbs = np.ascontiguousarray(mbs, dtype=abi.MBINFO_DTYPE)
            if mbs.shape != (self.mb_w * self.mb_h,):
                raise ValueError("mbs must hold one record per macroblock")
        self._ok(library().ia_set_reference(self._h, slot, abi.ptr(pic.y), abi.ptr(pic.cb), abi.ptr(pic.cr),
                                            abi.ptr(mbs)))

    def promote_reference(self, slot, deblock=True):
        """Makes the last analysed picture the reference in `slot` on the device: loop-filtered first
        when `deblock` (as the encoder keeps its references), then its half-sample planes built."""
        self._ok(library().ia_promote_reference(self._h, slot, 1 if deblock else 0))

    def read_reference(self, slot):
        """(Picture, macroblock records) of reference `slot`."""
        y = np.empty((self.ph, self.pw), dtype=np.uint8)
        cb = np.empty((self.ph // 2, self.pw // 2), dtype=np.uint8)
        cr = np.empty_like(cb)
        mbs = np.empty(self.mb_w * self.mb_h, dtype=abi.MBINFO_DTYPE)
        self._ok(library().ia_read_reference(self._h, slot, abi.ptr(y), abi.ptr(cb), abi.ptr(cr), abi.ptr(mbs)))
        return Picture(y, cb, cr), mbs

    def read_halfpel(self, slot):
        sw, sh = self.pw + 96, self.ph + 96
        out = np.empty((4, sh, sw), dtype=np.uint8)
        self._ok(library().ia_read_halfpel(self._h, slot, abi.ptr(out)))
        return out

    def upload(self, pic):
        pic = self._pic(pic)
        self._pic_keep = pic  # the copy is asynchronous: keep the host planes alive until the next sync
        self._ok(library().ia_upload_source(self._h, abi.ptr(pic.y), abi.ptr(pic.cb), abi.ptr(pic.cr)))

    def motion_search(self, ref_slots):
        refs = abi.int_array(list(ref_slots))
        self._ok(library().ia_motion_search(self._h, refs, len(ref_slots)))

    def read_motion(self, nref):
        out = np.empty((nref, self.mb_w * self.mb_h, abi.NPART), dtype=abi.FSRESULT_DTYPE)
        self._ok(library().ia_read_motion(self._h, abi.ptr(out)))
        return out

    def analyse_uploaded(self, qp, kind, ref_slots=()):
        """Enqueues the analysis of the uploaded picture (results stay on the device)."""
        refs = abi.int_array(list(ref_slots))
        self._ok(library().ia_analyse_uploaded(self._h, qp, kind, refs, len(ref_slots)))

    def sync(self):
        self._ok(library().ia_sync(self._h))

    def analyse(self, pic, qp, kind, ref_slots=()):
        """Analyses every macroblock of `pic` (`code_mb` across all slices) and reads the results back."""
        pic = self._pic(pic)
        n = self.mb_w * self.mb_h
        codes = np.zeros(n, dtype=abi.MBCODE_DTYPE)
        stats = np.zeros(n, dtype=abi.MBSTATS_DTYPE)
        mbs = np.zeros(n, dtype=abi.MBINFO_DTYPE)
        ry = np.empty_like(pic.y)
        rcb = np.empty_like(pic.cb)
        rcr = np.empty_like(pic.cr)
        refs = abi.int_array(list(ref_slots))
        self._ok(library().ia_analyse_picture(self._h, abi.ptr(pic.y), abi.ptr(pic.cb), abi.ptr(pic.cr), qp, kind,
                                              refs, len(ref_slots), abi.ptr(codes), abi.ptr(stats), abi.ptr(ry),
                                              abi.ptr(rcb), abi.ptr(rcr), abi.ptr(mbs)))
        return Analysis(codes, stats, Picture(ry, rcb, rcr), mbs, self.mb_w, self.mb_h, self.rows)

    def analyse_b(self, pic, qp, l0_slot, l1_slot, dist_l0, dist_l1):
        """Analyses `pic` as a B picture between the references in l0_slot (shown before it, dist_l0
        picture-order units earlier) and l1_slot (shown after it, dist_l1 units after the list 0 one)
        and reads the results back."""
        self.upload(pic)
        self.analyse_b_uploaded(qp, l0_slot, l1_slot, dist_l0, dist_l1)
        return self.read_analysis()

    def analyse_b_uploaded(self, qp, l0_slot, l1_slot, dist_l0, dist_l1):
        """Enqueues the B-picture analysis of the uploaded picture (results stay on the device)."""
        self._ok(library().ia_analyse_b(self._h, qp, l0_slot, l1_slot, dist_l0, dist_l1))

    def read_analysis(self):
        """The results of the last analysis of this context (e.g. after a batch step)."""
        n = self.mb_w * self.mb_h
        codes = np.zeros(n, dtype=abi.MBCODE_DTYPE)
        stats = np.zeros(n, dtype=abi.MBSTATS_DTYPE)
        mbs = np.zeros(n, dtype=abi.MBINFO_DTYPE)
        ry = np.empty((self.ph, self.pw), dtype=np.uint8)
        rcb = np.empty((self.ph // 2, self.pw // 2), dtype=np.uint8)
        rcr = np.empty_like(rcb)
        self._ok(library().ia_read_analysis(self._h, abi.ptr(codes), abi.ptr(stats), abi.ptr(ry), abi.ptr(rcb),
                                            abi.ptr(rcr), abi.ptr(mbs)))
        return Analysis(codes, stats, Picture(ry, rcb, rcr), mbs, self.mb_w, self.mb_h, self.rows)

    def load_proposal_weights(self, params):
        """Loads the learned proposal's CNN parameters (abi.CNN_PARAMS float32, see proposal.py)."""
        params = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._ok(library().ia_load_proposal_weights(self._h, abi.ptr(params), params.size))

    def proposal_features(self, ref_slots):
        """CNN input features [mb][81][8] and grid centres [mb][3][2] of the uploaded picture."""
        n = self.mb_w * self.mb_h
        feat = np.empty((n, abi.CNN_CELLS, abi.CNN_FEATURES), dtype=np.float32)
        ctr = np.empty((n, 3, 2), dtype=np.int16)
        refs = abi.int_array(list(ref_slots))
        self._ok(library().ia_proposal_features(self._h, refs, len(ref_slots), abi.ptr(feat), abi.ptr(ctr)))
        return feat, ctr

    def proposal_forward(self, feat):
        """The CNN's outputs [n][81][4] on given features (BF16 MFMA path)."""
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty((feat.shape[0], abi.CNN_CELLS, abi.CNN_OUTPUTS), dtype=np.float32)
        self._ok(library().ia_proposal_forward(self._h, abi.ptr(feat), feat.shape[0], abi.ptr(out)))
        return out

    def fallback_count(self):
        c = C.c_int()
        self._ok(library().ia_read_fallback_count(self._h, C.byref(c)))
        return c.value

    def set_fullsearch_variant(self, v):
        self._ok(library().ia_set_fullsearch_variant(self._h, v))


def fullsearch_variants():
    lib = library()
    return [lib.ia_fullsearch_variant_name(i).decode() for i in range(lib.ia_fullsearch_variants())]


class Batch:
    """Streams of any sizes (Analysers) through one set of launches per frame step; with use_graph the
    step is one captured hipGraph, replayed every frame."""

    def __init__(self, analysers, use_graph=True):
        self.analysers = list(analysers)
        n = len(self.analysers)
        arr = (C.c_void_p * n)(*[a._h.value for a in self.analysers])
        h = C.c_void_p()
        rc = library().ia_batch_create(C.byref(h), arr, n, int(use_graph))
        if rc != abi.IA_OK:
            if rc == abi.IA_EINVAL:
                raise ValueError("ia_batch_create: invalid arguments")
            raise AnalysisError(f"ia_batch_create failed ({rc})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            library().ia_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, frames):
        """frames[i] for analyser i: dict(qp, kind, refs=[slots newest first], promote=slot or -1,
        deblock=True (the promoted reference loop-filtered first, as the encoder keeps it),
        planes=(y, cb, cr) device addresses or None for the analyser's uploaded source). A B picture
        (kind abi.PIC_B) takes refs=[list 0 slot, list 1 slot] and dist=(dist_l0, dist_l1)."""
        if len(frames) != len(self.analysers):
            raise ValueError("one frame per stream")
        arr = (abi.Frame * len(frames))()
        for f, d in zip(arr, frames):
            refs = list(d.get("refs", ()))
            f.qp, f.kind, f.nref = d["qp"], d["kind"], len(refs)
            for i, r in enumerate(refs):
                f.ref_slots[i] = r
            f.promote_slot = d.get("promote", -1)
            f.promote_deblock = 1 if d.get("deblock", True) else 0
            f.dist_l0, f.dist_l1 = d.get("dist", (0, 0))
            planes = d.get("planes")
            if planes is not None:
                f.y, f.cb, f.cr = planes
        rc = library().ia_batch_step(self._h, arr)
        if rc != abi.IA_OK:
            msg = library().ia_batch_last_error(self._h).decode(errors="replace")
            if rc == abi.IA_EINVAL:
                raise ValueError(msg)
            raise AnalysisError(f"error {rc}: {msg}")

    def fallback_count(self):
        c = C.c_int()
        rc = library().ia_batch_fallback_count(self._h, C.byref(c))
        if rc != abi.IA_OK:
            raise AnalysisError(library().ia_batch_last_error(self._h).decode(errors="replace"))
        return c.value

    def sync(self):
        rc = library().ia_batch_sync(self._h)
        if rc != abi.IA_OK:
            raise AnalysisError(library().ia_batch_last_error(self._h).decode(errors="replace"))
