"""ctypes mirror of include/ironav_h264a.h (structs, constants and function prototypes)."""
import ctypes as C

import numpy as np

IA_OK, IA_EINVAL, IA_EDEVICE, IA_ENOMEM, IA_ESTATE, IA_EUNSUPPORTED = 0, -1, -2, -3, -4, -5
KIND_NONE, KIND_I4X4, KIND_I8X8, KIND_I16X16, KIND_IPCM, KIND_INTER = range(6)
SEARCH_DIAMOND, SEARCH_FULL, SEARCH_LEARNED = 0, 1, 2
PIC_IDR, PIC_P, PIC_B = 0, 1, 2
NPART = 41
CNN_PARAMS, CNN_CELLS, CNN_FEATURES, CNN_OUTPUTS = 21988, 81, 8, 4

# (x, y, w, h) in 4x4 blocks of the 41 searched partitions, numbered as IA_NPART documents.
PARTITIONS = (
    [(0, 0, 4, 4), (0, 0, 4, 2), (0, 2, 4, 2), (0, 0, 2, 4), (2, 0, 2, 4)]
    + [((b & 1) * 2, (b >> 1) * 2, 2, 2) for b in range(4)]
    + [((b & 1) * 2, (b >> 1) * 2 + h, 2, 1) for b in range(4) for h in range(2)]
    + [((b & 1) * 2 + h, (b >> 1) * 2, 1, 2) for b in range(4) for h in range(2)]
    + [((b & 1) * 2 + (j & 1), (b >> 1) * 2 + (j >> 1), 1, 1) for b in range(4) for j in range(4)]
)


class Config(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32), ("cabac", C.c_int32),
        ("transform8", C.c_int32), ("sub8x8", C.c_int32), ("fast", C.c_int32), ("nrefs", C.c_int32),
        ("slices", C.c_int32), ("search_mode", C.c_int32), ("search_range", C.c_int32),
        ("fallback_ratio", C.c_float),
    ]


class MbInfo(C.Structure):
    _fields_ = [
        ("slice", C.c_uint16), ("kind", C.c_uint8), ("qp", C.c_int8), ("cbp", C.c_uint8), ("skip", C.c_uint8),
        ("t8x8", C.c_uint8), ("reserved", C.c_uint8), ("nnz", C.c_uint8 * 16), ("mv", (C.c_int16 * 2) * 16),
        ("ref_idx", C.c_int8 * 4), ("ref_id", C.c_int32 * 4), ("mv1", (C.c_int16 * 2) * 16),
        ("ref_idx1", C.c_int8 * 4), ("ref_id1", C.c_int32 * 4), ("bdirect", C.c_uint8), ("direct8", C.c_uint8),
        ("reserved1", C.c_uint8 * 2),
    ]


class MbCode(C.Structure):
    _fields_ = [
        ("kind", C.c_uint8), ("skip", C.c_uint8), ("i16_mode", C.c_uint8), ("chroma_mode", C.c_uint8),
        ("part", C.c_uint8), ("t8x8", C.c_uint8), ("cbp", C.c_uint8), ("reserved", C.c_uint8),
        ("modes", C.c_uint8 * 16), ("refs", C.c_uint8 * 4), ("sub", C.c_uint8 * 4), ("nnz", C.c_uint8 * 16),
        ("cnnz", (C.c_uint8 * 4) * 2), ("mvs", (C.c_int16 * 2) * 16), ("mvds", (C.c_int16 * 2) * 16),
        ("luma", (C.c_int16 * 16) * 16), ("luma8", (C.c_int16 * 64) * 4), ("dc", C.c_int16 * 16),
        ("cdc", (C.c_int16 * 4) * 2), ("cac", ((C.c_int16 * 16) * 4) * 2), ("bdir", C.c_uint8 * 4),
        ("buse", C.c_uint8 * 4), ("bmv", ((C.c_int16 * 2) * 16) * 2), ("bmvd", ((C.c_int16 * 2) * 4) * 2),
    ]


class MbStats(C.Structure):
    _fields_ = [
        ("skip_sad", C.c_int32), ("i16_cost", C.c_int32), ("inter_cost", C.c_int32), ("inter_part", C.c_int32),
        ("rd_cost", C.c_double), ("bits", C.c_int32),
    ]


class FsResult(C.Structure):
    _fields_ = [("mv", C.c_int16 * 2), ("sad", C.c_uint32)]


class Frame(C.Structure):
    """ia_frame: one stream's picture in a batch step."""
    _fields_ = [
        ("qp", C.c_int), ("kind", C.c_int), ("nref", C.c_int), ("ref_slots", C.c_int * 16),
        ("promote_slot", C.c_int), ("y", C.c_void_p), ("cb", C.c_void_p), ("cr", C.c_void_p),
        ("promote_deblock", C.c_int), ("dist_l0", C.c_int), ("dist_l1", C.c_int),
    ]


MBINFO_DTYPE = np.dtype(MbInfo)
MBCODE_DTYPE = np.dtype(MbCode)
MBSTATS_DTYPE = np.dtype(MbStats)
FSRESULT_DTYPE = np.dtype(FsResult)

u8p = C.POINTER(C.c_uint8)
vp = C.c_void_p
ip = C.POINTER(C.c_int)

# name -> (restype, argtypes): every entry point include/ironav_h264a.h declares
PROTOTYPES = {
    "ia_create": (C.c_int, [C.POINTER(vp), C.POINTER(Config)]),
    "ia_destroy": (None, [vp]),
    "ia_last_error": (C.c_char_p, [vp]),
    "ia_set_stream": (C.c_int, [vp, vp]),
    "ia_set_reference": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
    "ia_read_halfpel": (C.c_int, [vp, C.c_int, vp]),
    "ia_promote_reference": (C.c_int, [vp, C.c_int, C.c_int]),
    "ia_read_reference": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
    "ia_upload_source": (C.c_int, [vp, vp, vp, vp]),
    "ia_motion_search": (C.c_int, [vp, ip, C.c_int]),
    "ia_read_motion": (C.c_int, [vp, vp]),
    "ia_analyse_uploaded": (C.c_int, [vp, C.c_int, C.c_int, ip, C.c_int]),
    "ia_analyse_b": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ia_read_analysis": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "ia_analyse_picture": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, ip, C.c_int, vp, vp, vp, vp, vp, vp]),
    "ia_sync": (C.c_int, [vp]),
    "ia_load_proposal_weights": (C.c_int, [vp, vp, C.c_int]),
    "ia_proposal_features": (C.c_int, [vp, ip, C.c_int, vp, vp]),
    "ia_proposal_forward": (C.c_int, [vp, vp, C.c_int, vp]),
    "ia_read_fallback_count": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ia_fullsearch_variants": (C.c_int, []),
    "ia_fullsearch_variant_name": (C.c_char_p, [C.c_int]),
    "ia_set_fullsearch_variant": (C.c_int, [vp, C.c_int]),
    "ia_batch_create": (C.c_int, [C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int]),
    "ia_batch_destroy": (None, [vp]),
    "ia_batch_last_error": (C.c_char_p, [vp]),
    "ia_batch_step": (C.c_int, [vp, C.POINTER(Frame)]),
    "ia_batch_sync": (C.c_int, [vp]),
    "ia_batch_fallback_count": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "ia_batch_workers": (C.c_int, []),
    "ia_deblock_picture_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp, vp]),
}


def bind(lib):
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def int_array(vals):
    arr = (C.c_int * max(1, len(vals)))(*vals)
    return arr
