"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the GPU path. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it; the product (ironav_amd/) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ironav_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i = C.c_void_p, C.c_int
        sig = {
            "or_halfpel_build": (None, [vp, i, i, i, vp]),
            "or_halfpel_predict": (None, [vp, i, i, i, i, i, i, i, i, vp]),
            "or_luma_mc_ref": (None, [vp, i, i, i, i, i, i, i, i, i, vp, i]),
            "or_chroma_mc": (None, [vp, i, i, i, i, i, i, i, i, vp]),
            "or_fdct4": (None, [vp, vp]),
            "or_fdct8": (None, [vp, vp]),
            "or_idct4_add": (None, [vp, vp, i, i]),
            "or_idct8_add": (None, [vp, vp, i, i]),
            "or_hadamard4": (None, [vp, vp]),
            "or_quant": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, i]),
            "or_quant8": (None, [vp, vp, i, i, vp, vp]),
            "or_satd4": (C.c_int32, [vp]),
            "or_decimate_score": (C.c_int32, [vp]),
            "or_decimate_score8": (C.c_int32, [vp]),
            "or_ue_bits": (C.c_uint32, [C.c_uint32]),
            "or_se_bits": (C.c_uint32, [C.c_int32]),
            "or_lambda": (C.c_int32, [i]),
            "or_entropy_table": (None, [vp]),
            "or_predict_nxn": (None, [i, i, vp, vp, C.c_int32, i, i, i, vp, i]),
            "or_filter_8x8_edges": (None, [vp, vp, vp, i, i, i]),
            "or_predict_16x16": (None, [i, vp, vp, C.c_int32, i, i, vp, i]),
            "or_predict_chroma": (None, [i, vp, vp, C.c_int32, i, i, vp, i, i]),
            "or_fullsearch": (None, [vp, vp, i, i, i, i, vp]),
            "or_abi_sizes": (None, [vp]),
            "or_deblock_picture": (None, [vp, vp, vp, vp, i, i]),
            "or_analyse_picture": (i, [C.POINTER(abi.Config), vp, vp, vp, i, i, vp, vp, vp, vp, i, vp, vp, vp, vp, vp,
                                       vp, vp]),
            "or_analyse_picture_mt": (i, [C.POINTER(abi.Config), vp, vp, vp, i, i, vp, vp, vp, vp, i, vp, vp, vp, vp,
                                          vp, vp, vp, i]),
            "or_analyse_picture_b": (i, [C.POINTER(abi.Config), vp, vp, vp, i, vp, vp, vp, vp, i, i, vp, vp, vp, vp,
                                         vp, vp, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


P = abi.ptr


def halfpel_build(luma16, bd=8):
    ph, pw = luma16.shape
    luma16 = np.ascontiguousarray(luma16, dtype=np.uint16)
    out = np.empty((4, ph + 96, pw + 96), dtype=np.uint16)
    lib().or_halfpel_build(P(luma16), pw, ph, bd, P(out))
    return out


def deblock_picture(y, cb, cr, mbs, mb_w, mb_h):
    """The encoder's loop filter (h264/deblock.rs:608-695) on copies of padded 8-bit planes y
    (mb_h*16 x mb_w*16), cb and cr (mb_h*8 x mb_w*8) with the picture's macroblock records; returns the
    filtered (y, cb, cr)."""
    y = np.ascontiguousarray(y, dtype=np.uint8).copy()
    cb = np.ascontiguousarray(cb, dtype=np.uint8).copy()
    cr = np.ascontiguousarray(cr, dtype=np.uint8).copy()
    assert y.shape == (mb_h * 16, mb_w * 16) and cb.shape == cr.shape == (mb_h * 8, mb_w * 8)
    mbs = np.ascontiguousarray(mbs, dtype=abi.MBINFO_DTYPE)
    assert mbs.size == mb_w * mb_h
    lib().or_deblock_picture(P(y), P(cb), P(cr), P(mbs), mb_w, mb_h)
    return y, cb, cr


def halfpel_predict(planes, pw, ph, x, y, mv, w, h):
    out = np.empty(w * h, dtype=np.uint16)
    lib().or_halfpel_predict(P(planes), pw, ph, x, y, int(mv[0]), int(mv[1]), w, h, P(out))
    return out


def luma_mc_ref(luma16, x, y, mv, w, h, maxv):
    ph, pw = luma16.shape
    luma16 = np.ascontiguousarray(luma16, dtype=np.uint16)
    out = np.empty(w * h, dtype=np.uint16)
    lib().or_luma_mc_ref(P(luma16), pw, pw, ph, x, y, int(mv[0]), int(mv[1]), w, h, P(out), maxv)
    return out


def fullsearch(src_y, ref_ys, mb_w, mb_h, rng):
    refs = [np.ascontiguousarray(r, dtype=np.uint8) for r in ref_ys]
    arr = (C.c_void_p * len(refs))(*[r.ctypes.data for r in refs])
    out = np.empty((len(refs), mb_w * mb_h, abi.NPART), dtype=abi.FSRESULT_DTYPE)
    lib().or_fullsearch(P(np.ascontiguousarray(src_y)), C.cast(arr, C.c_void_p), len(refs), mb_w, mb_h, rng, P(out))
    return out


def analyse_picture(cfg, pic, qp, kind, refs=(), fs=None, threads=1):
    """refs: sequence of (Picture, mbinfo array or None), list-0 order. Returns
    (codes, stats, (ry, rcb, rcr), mbs). threads > 1 codes the slices on that many host threads."""
    mb_w, mb_h = (cfg.width + 15) // 16, (cfg.height + 15) // 16
    n = mb_w * mb_h
    codes = np.zeros(n, dtype=abi.MBCODE_DTYPE)
    stats = np.zeros(n, dtype=abi.MBSTATS_DTYPE)
    mbs = np.zeros(n, dtype=abi.MBINFO_DTYPE)
    ry, rcb, rcr = np.empty_like(pic.y), np.empty_like(pic.cb), np.empty_like(pic.cr)
    keep = []

    def arr(items):
        a = (C.c_void_p * max(1, len(items)))(*[x.ctypes.data if x is not None else None for x in items])
        keep.append(a)
        return C.cast(a, C.c_void_p)

    ref_mbs = [None if m is None else np.ascontiguousarray(m, dtype=abi.MBINFO_DTYPE) for _, m in refs]
    rc = lib().or_analyse_picture_mt(C.byref(cfg), P(pic.y), P(pic.cb), P(pic.cr), qp, kind,
                                     arr([r.y for r, _ in refs]), arr([r.cb for r, _ in refs]),
                                     arr([r.cr for r, _ in refs]), arr(ref_mbs) if refs else None, len(refs),
                                     P(fs) if fs is not None else None, P(codes), P(stats), P(ry), P(rcb), P(rcr),
                                     P(mbs), threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return codes, stats, (ry, rcb, rcr), mbs


def analyse_picture_b(cfg, pic, qp, l0, l1, tb, td, threads=1):
    """A B picture's analysis: l0 and l1 are (Picture, mbinfo array or None) of the list 0 and list 1
    pictures, tb/td the picture-order distances from the list 0 picture to this one and to the list 1
    picture. Returns (codes, stats, (ry, rcb, rcr), mbs)."""
    mb_w, mb_h = (cfg.width + 15) // 16, (cfg.height + 15) // 16
    n = mb_w * mb_h
    codes = np.zeros(n, dtype=abi.MBCODE_DTYPE)
    stats = np.zeros(n, dtype=abi.MBSTATS_DTYPE)
    mbs = np.zeros(n, dtype=abi.MBINFO_DTYPE)
    ry, rcb, rcr = np.empty_like(pic.y), np.empty_like(pic.cb), np.empty_like(pic.cr)
    refs = (l0, l1)
    ref_mbs = [None if m is None else np.ascontiguousarray(m, dtype=abi.MBINFO_DTYPE) for _, m in refs]

    def arr(items):
        return (C.c_void_p * 2)(*[x.ctypes.data if x is not None else None for x in items])

    ay, acb, acr, am = arr([r.y for r, _ in refs]), arr([r.cb for r, _ in refs]), arr([r.cr for r, _ in refs]), arr(ref_mbs)
    rc = lib().or_analyse_picture_b(C.byref(cfg), P(pic.y), P(pic.cb), P(pic.cr), qp, C.cast(ay, C.c_void_p),
                                    C.cast(acb, C.c_void_p), C.cast(acr, C.c_void_p), C.cast(am, C.c_void_p), tb, td,
                                    P(codes), P(stats), P(ry), P(rcb), P(rcr), P(mbs), threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return codes, stats, (ry, rcb, rcr), mbs
