"""Host side around the GPU analysis: the encoder driver's picture-type decisions and the bitstream
back end (include/ironav_h264w.h, ironav_amd/libironav_h264w.so).

`Stream` writes each analysed picture (the `ia_mbcode`s of ia_read_analysis) as an Annex B access
unit, as the reference encoder's `code_picture` writes its slices with `write_mb_cabac`
(crates/ironav-codec/src/h264/enc/mod.rs:539-652, enc/frame.rs:2020-2279).

`Gop` restates the picture-type rule of nervosys/IronAV's `H264Encoder::encode`
(crates/ironav-codec/src/h264/enc/mod.rs:700-717): IDR for the first picture, after `keyint`
pictures, when forced, or at a scene cut measured by `change_threads`
(crates/ironav-codec/src/change.rs:57-95); P otherwise (the IPPP coding of the workloads).
"""
import ctypes as C
import os

import numpy as np

from . import abi

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libironav_h264w.so")

IW_OK, IW_EINVAL, IW_ENOMEM, IW_EUNSUPPORTED, IW_ESPACE = 0, -1, -3, -5, -6


class Params(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fps_num", C.c_int32), ("fps_den", C.c_int32),
                ("cabac", C.c_int32), ("transform8", C.c_int32), ("nrefs", C.c_int32), ("slices", C.c_int32),
                ("bframes", C.c_int32)]


def library():
    """The loaded host library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ironav_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i, i64p = C.c_void_p, C.c_int, C.POINTER(C.c_int64)
        sig = {
            "iw_lowres": (i, [vp, i, i, vp]),
            "iw_change": (i, [vp, vp, i, i, i, i64p, i64p]),
            "iw_gop_create": (i, [C.POINTER(vp), i, i, C.c_uint64, i]),
            "iw_gop_destroy": (None, [vp]),
            "iw_gop_next": (i, [vp, vp, i, C.POINTER(C.c_int), i64p, i64p]),
            "iw_stream_create": (i, [C.POINTER(vp), C.POINTER(Params)]),
            "iw_stream_destroy": (None, [vp]),
            "iw_write_picture": (i, [vp, i, i, i, vp, C.POINTER(C.c_size_t)]),
            "iw_write_picture_at": (i, [vp, i, i, i, C.c_int64, vp, C.POINTER(C.c_size_t)]),
            "iw_read_picture": (i, [vp, vp, C.c_size_t]),
            "iw_level": (i, [i, i, i, i]),
            "iw_put_nal": (i, [i, i, vp, C.c_size_t, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
            "iw_cabac_encode_bins": (i, [vp, vp, vp, i, i, i, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


def _check(rc):
    if rc == IW_OK:
        return
    if rc == IW_EINVAL:
        raise ValueError("invalid argument")
    if rc == IW_EUNSUPPORTED:
        raise NotImplementedError("the bitstream back end writes CABAC streams only")
    raise RuntimeError(f"ironav_h264w error {rc}")


def lowres(y):
    """4x4-block means (rounded down) of a padded luma plane: (mb_h*4, mb_w*4) int32."""
    y = np.ascontiguousarray(y, dtype=np.uint8)
    mb_h, mb_w = y.shape[0] // 16, y.shape[1] // 16
    out = np.empty((mb_h * 4, mb_w * 4), dtype=np.int32)
    _check(library().iw_lowres(abi.ptr(y), mb_w, mb_h, abi.ptr(out)))
    return out


def change(prev, cur, threads=1):
    """(inter, intra) of two low-resolution pictures (change.rs:57-89)."""
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    cur = np.ascontiguousarray(cur, dtype=np.int32)
    if prev.shape != cur.shape:
        raise ValueError("low-resolution pictures of different sizes")
    a, b = C.c_int64(), C.c_int64()
    _check(library().iw_change(abi.ptr(prev), abi.ptr(cur), cur.shape[1], cur.shape[0], threads, C.byref(a),
                               C.byref(b)))
    return a.value, b.value


class Gop:
    """Picture types of one encoder's input pictures (IPPP with IDR pictures at scene cuts)."""

    def __init__(self, width, height, keyint=0, threads=1):
        h = C.c_void_p()
        _check(library().iw_gop_create(C.byref(h), width, height, keyint, threads))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            library().iw_gop_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def next(self, y, force_idr=False):
        """abi.PIC_IDR or abi.PIC_P for the next picture (padded luma y)."""
        y = np.ascontiguousarray(y, dtype=np.uint8)
        k = C.c_int()
        _check(library().iw_gop_next(self._h, abi.ptr(y), int(bool(force_idr)), C.byref(k), None, None))
        return abi.PIC_IDR if k.value == 0 else abi.PIC_P


def coding_order(idr, bframes=2):
    """The order `H264Encoder::encode` codes pictures in (enc/mod.rs:700-731, code_pending :675-686):
    idr[d] says whether display picture d is an IDR picture (Gop.next); with bframes > 0, the pictures
    between two P pictures wait and are coded as B pictures after the P picture that follows them; the
    last waiting picture before an IDR picture, or at the end, is coded as P instead. Returns
    [(display, kind)] in coding order, kind abi.PIC_IDR / PIC_P / PIC_B."""
    out, pending = [], []

    def code_pending():
        if pending:
            out.append((pending.pop(), abi.PIC_P))
        out.extend((d, abi.PIC_B) for d in pending)
        pending.clear()

    for d, is_idr in enumerate(idr):
        if is_idr or len(pending) == bframes:
            if is_idr:
                code_pending()
            out.append((d, abi.PIC_IDR if is_idr else abi.PIC_P))
            out.extend((p, abi.PIC_B) for p in pending)
            pending.clear()
        else:
            pending.append(d)
    code_pending()
    return out


def level(width, height, fps_num=0, fps_den=0):
    """level_idc the SPS carries for this size and rate."""
    return int(library().iw_level(width, height, fps_num, fps_den))


def put_nal(ref_idc, nal_type, rbsp):
    """A NAL unit with its start code and emulation prevention bytes."""
    rbsp = bytes(rbsp)
    n = C.c_size_t()
    _check(library().iw_put_nal(ref_idc, nal_type, rbsp, len(rbsp), None, 0, C.byref(n)))
    out = C.create_string_buffer(n.value)
    _check(library().iw_put_nal(ref_idc, nal_type, rbsp, len(rbsp), out, n.value, C.byref(n)))
    return out.raw[:n.value]


def cabac_encode_bins(kinds, ctx, vals, intra, qp):
    """The CABAC coder alone: bins (kind 0 decision, 1 bypass, 2 terminate 0) then a terminate 1."""
    kinds = np.ascontiguousarray(kinds, dtype=np.uint8)
    ctx = np.ascontiguousarray(ctx, dtype=np.uint16)
    vals = np.ascontiguousarray(vals, dtype=np.uint8)
    n = C.c_size_t()
    L = library()
    _check(L.iw_cabac_encode_bins(abi.ptr(kinds), abi.ptr(ctx), abi.ptr(vals), len(kinds), int(intra), qp, None, 0,
                                  C.byref(n)))
    out = C.create_string_buffer(max(1, n.value))
    _check(L.iw_cabac_encode_bins(abi.ptr(kinds), abi.ptr(ctx), abi.ptr(vals), len(kinds), int(intra), qp, out,
                                  n.value, C.byref(n)))
    return out.raw[:n.value]


class Stream:
    """One output stream: each analysed picture written as an Annex B access unit. With bframes > 0
    pictures are written in coding order with their display index (write_at), B pictures included."""

    def __init__(self, width, height, *, nrefs=3, slices=1, transform8=True, cabac=True, fps=(0, 0), bframes=0):
        p = Params(width, height, fps[0], fps[1], int(cabac), int(transform8), nrefs, slices, bframes)
        h = C.c_void_p()
        _check(library().iw_stream_create(C.byref(h), C.byref(p)))
        self._h = h
        self.mb_w, self.mb_h = (width + 15) // 16, (height + 15) // 16

    def close(self):
        if getattr(self, "_h", None):
            library().iw_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def write(self, kind, qp, codes, nref=0):
        """The access unit of a picture of `kind` (abi.PIC_IDR / abi.PIC_P against `nref` references)."""
        codes = np.ascontiguousarray(codes, dtype=abi.MBCODE_DTYPE)
        if codes.shape != (self.mb_w * self.mb_h,):
            raise ValueError("one code per macroblock")
        n = C.c_size_t()
        L = library()
        _check(L.iw_write_picture(self._h, 0 if kind == abi.PIC_IDR else 1, qp, nref, abi.ptr(codes), C.byref(n)))
        out = C.create_string_buffer(max(1, n.value))
        _check(L.iw_read_picture(self._h, out, n.value))
        return out.raw[:n.value]

    def write_at(self, kind, qp, codes, display, nref=0):
        """The access unit of a picture of `kind` (abi.PIC_IDR, PIC_P or PIC_B) shown at `display`."""
        codes = np.ascontiguousarray(codes, dtype=abi.MBCODE_DTYPE)
        if codes.shape != (self.mb_w * self.mb_h,):
            raise ValueError("one code per macroblock")
        n = C.c_size_t()
        L = library()
        _check(L.iw_write_picture_at(self._h, int(kind), qp, nref, int(display), abi.ptr(codes), C.byref(n)))
        out = C.create_string_buffer(max(1, n.value))
        _check(L.iw_read_picture(self._h, out, n.value))
        return out.raw[:n.value]
