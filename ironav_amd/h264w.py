"""Host side around the GPU analysis: the encoder driver's picture-type decisions and the bitstream
back end (include/ironav_h264w.h, ironav_amd/libironav_h264w.so).

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

IW_OK, IW_EINVAL, IW_ENOMEM, IW_ESPACE = 0, -1, -3, -6


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
