"""ctypes bindings of the oracle's H.264 decoder (oracle/decoder.c) and a minimal MP4 sample reader.

TEST INFRASTRUCTURE ONLY: the verifier of bitstreams. Only tests/ and bench.py's measurement of
decoded rates use it; the product (ironav_amd/) never does.
"""
import ctypes as C
import struct

import numpy as np

from ironav_amd import abi

from . import pyoracle


def _lib():
    L = pyoracle.lib()
    if not getattr(L, "_od_bound", False):
        vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
        sig = {
            "od_create": (vp, []),
            "od_destroy": (None, [vp]),
            "od_error": (C.c_char_p, [vp]),
            "od_decode_nal": (i, [vp, vp, sz]),
            "od_decode_annexb": (i, [vp, vp, sz]),
            "od_flush": (i, [vp]),
            "od_frame_count": (i, [vp]),
            "od_frame_size": (i, [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
            "od_frame_read": (i, [vp, i, vp, vp, vp]),
            "od_frame_read_full": (i, [vp, i, vp, vp, vp, vp]),
            "od_cabac_decode_bins": (i, [vp, sz, vp, vp, i, i, i, vp, C.POINTER(i), C.POINTER(sz)]),
            "od_stats": (i, [vp, vp, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        L._od_bound = True
    return L


class DecodeError(RuntimeError):
    pass


class Frame:
    def __init__(self, y, cb, cr, full=None, mbs=None):
        self.y, self.cb, self.cr = y, cb, cr  # cropped planes
        self.full = full                      # (y, cb, cr) whole macroblock planes, loop-filtered
        self.mbs = mbs                        # ia_mbinfo records


class Decoder:
    def __init__(self):
        self.L = _lib()
        self.h = self.L.od_create()
        if not self.h:
            raise MemoryError("od_create")

    def close(self):
        if self.h:
            self.L.od_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _ok(self, rc):
        if rc != 0:
            raise DecodeError(f"decoder error {rc}: {self.L.od_error(self.h).decode(errors='replace')}")

    def nal(self, data):
        b = bytes(data)
        self._ok(self.L.od_decode_nal(self.h, b, len(b)))

    def annexb(self, data):
        b = bytes(data)
        self._ok(self.L.od_decode_annexb(self.h, b, len(b)))

    def stats(self):
        """Counts of the syntax the decoded pictures exercised (decoder.h, od_stats)."""
        out = np.zeros(64, np.int64)
        n = self.L.od_stats(self.h, out.ctypes.data, 64)
        out = out[:n]
        return {"i4_modes": out[0:9], "i8_modes": out[9:18], "i16_modes": out[18:22], "chroma_modes": out[22:26],
                "partitions": out[26:30], "sub_partitions": out[30:34], "skip": int(out[34]), "pcm": int(out[35]),
                "inter_t8": int(out[36]), "ref_gt0": int(out[37]), "b_skip": int(out[38]), "b_direct16": int(out[39]),
                "b16": out[40:43], "b_partitions": out[43:45], "b8x8": int(out[45]), "b_sub": out[46:59],
                "b_temporal": int(out[59]), "b_implicit": int(out[60]), "b_explicit": int(out[61])}

    def frames(self):
        self._ok(self.L.od_flush(self.h))
        out = []
        for k in range(self.L.od_frame_count(self.h)):
            w, h, mw, mh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            self._ok(self.L.od_frame_size(self.h, k, C.byref(w), C.byref(h), C.byref(mw), C.byref(mh)))
            W, H = w.value, h.value
            y = np.empty((H, W), np.uint8)
            cb = np.empty(((H + 1) // 2, (W + 1) // 2), np.uint8)
            cr = np.empty_like(cb)
            self._ok(self.L.od_frame_read(self.h, k, y.ctypes.data, cb.ctypes.data, cr.ctypes.data))
            fy = np.empty((mh.value * 16, mw.value * 16), np.uint8)
            fcb = np.empty((mh.value * 8, mw.value * 8), np.uint8)
            fcr = np.empty_like(fcb)
            mbs = np.zeros(mw.value * mh.value, dtype=abi.MBINFO_DTYPE)
            self._ok(self.L.od_frame_read_full(self.h, k, fy.ctypes.data, fcb.ctypes.data, fcr.ctypes.data,
                                               mbs.ctypes.data))
            out.append(Frame(y, cb, cr, (fy, fcb, fcr), mbs))
        return out


def cabac_decode_bins(data, kinds, ctx, intra, qp):
    """(values, final terminate bin, bytes read) of the bins `kinds`/`ctx` decoded from `data`."""
    L = _lib()
    kinds = np.ascontiguousarray(kinds, dtype=np.uint8)
    ctx = np.ascontiguousarray(ctx, dtype=np.uint16)
    vals = np.zeros(len(kinds), np.uint8)
    fin, nread = C.c_int(), C.c_size_t()
    b = bytes(data)
    rc = L.od_cabac_decode_bins(b, len(b), kinds.ctypes.data, ctx.ctypes.data, len(kinds), int(intra), qp,
                                vals.ctypes.data, C.byref(fin), C.byref(nread))
    if rc != 0:
        raise DecodeError(f"od_cabac_decode_bins: {rc}")
    return vals, fin.value, nread.value


def decode_annexb(data):
    d = Decoder()
    try:
        d.annexb(data)
        return d.frames()
    finally:
        d.close()


# ---- MP4 (ISO/IEC 14496-12/15): the avcC parameter sets and the length-prefixed samples ----

def _boxes(b, off, end):
    while off + 8 <= end:
        size, typ = struct.unpack(">I4s", b[off:off + 8])
        hdr = 8
        if size == 1:
            size = struct.unpack(">Q", b[off + 8:off + 16])[0]
            hdr = 16
        elif size == 0:
            size = end - off
        yield typ.decode("latin1"), off + hdr, off + size
        off += size


def _find(b, off, end, path):
    for t, s, e in _boxes(b, off, end):
        if t == path[0]:
            if len(path) == 1:
                return s, e
            r = _find(b, s, e, path[1:])
            if r:
                return r
    return None


def mp4_nals(path):
    """The NAL units of the first video track: the SPS/PPS of avcC, then each sample's NAL units."""
    b = open(path, "rb").read()
    stbl = _find(b, 0, len(b), ["moov", "trak", "mdia", "minf", "stbl"])
    if not stbl:
        raise ValueError("no sample table")
    s0, e0 = stbl
    boxes = {t: (s, e) for t, s, e in _boxes(b, s0, e0)}
    ss, se = boxes["stsd"]
    nal_len, nals = 4, []
    for t, s, e in _boxes(b, ss + 8, se):
        if t in ("avc1", "avc3"):
            for t2, s2, e2 in _boxes(b, s + 78, e):
                if t2 == "avcC":
                    a = b[s2:e2]
                    nal_len = (a[4] & 3) + 1
                    p = 6
                    for _ in range(a[5] & 31):
                        n = struct.unpack(">H", a[p:p + 2])[0]
                        nals.append(a[p + 2:p + 2 + n])
                        p += 2 + n
                    for _ in range(a[p]):
                        n = struct.unpack(">H", a[p + 1:p + 3])[0]
                        nals.append(a[p + 3:p + 3 + n])
                        p += 2 + n
    # sample sizes, chunk offsets and the sample-to-chunk map
    s, e = boxes["stsz"]
    fixed, count = struct.unpack(">II", b[s + 4:s + 12])
    sizes = [fixed] * count if fixed else list(struct.unpack(f">{count}I", b[s + 12:s + 12 + 4 * count]))
    if "stco" in boxes:
        s, e = boxes["stco"]
        n = struct.unpack(">I", b[s + 4:s + 8])[0]
        offs = list(struct.unpack(f">{n}I", b[s + 8:s + 8 + 4 * n]))
    else:
        s, e = boxes["co64"]
        n = struct.unpack(">I", b[s + 4:s + 8])[0]
        offs = list(struct.unpack(f">{n}Q", b[s + 8:s + 8 + 8 * n]))
    s, e = boxes["stsc"]
    n = struct.unpack(">I", b[s + 4:s + 8])[0]
    runs = [struct.unpack(">III", b[s + 8 + 12 * k:s + 20 + 12 * k]) for k in range(n)]
    k = 0
    for ci, off in enumerate(offs):
        chunk = ci + 1
        per = [r for r in runs if r[0] <= chunk][-1][1]
        for _ in range(per):
            if k >= len(sizes):
                break
            sample = b[off:off + sizes[k]]
            off += sizes[k]
            k += 1
            p = 0
            while p + nal_len <= len(sample):
                n = int.from_bytes(sample[p:p + nal_len], "big")
                nals.append(sample[p + nal_len:p + nal_len + n])
                p += nal_len + n
    return nals


def decode_mp4(path, with_stats=False):
    d = Decoder()
    try:
        for n in mp4_nals(path):
            d.nal(n)
        frames = d.frames()
        return (frames, d.stats()) if with_stats else frames
    finally:
        d.close()
