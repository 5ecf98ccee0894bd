"""The oracle's H.264 decoder (oracle/decoder.c) against the reference decoder's own fixtures.

Each stream is a libx264 file the reference's engine tests decode and compare with FFmpeg's decoded
pictures (crates/ironav-engine/tests/e2e.rs:1991-2036); copies are committed under
tests/golden/h264_decoder/ (collect.py). The subset here is every progressive 8-bit 4:2:0 CABAC stream
with I, P and B slices: intra 4x4/8x8/16x16 and I_PCM, the 8x8 transform, scaling matrices, cropping,
several slices, every partition size, four references, explicit weighted prediction, constrained intra
prediction; B_Skip, B_Direct_16x16 and B_8x8 with spatial and with temporal direct prediction,
implicit bi-prediction weights, B pictures used as references (a pyramid). Every picture must equal
FFmpeg's byte for byte.

Decoding them exercises the oracle's reconstruction primitives, which the decoder calls unchanged:
the quarter-sample luma and eighth-sample chroma interpolation (or_luma_mc_ref, or_chroma_mc), all
nine 4x4/8x8 intra modes with the 8x8 reference filter, the 16x16 and chroma modes, the inverse
transforms and the loop filter (or_deblock_picture_params); and the CABAC context initialisation
tables with the arithmetic decoder the encoder's rate model mirrors.
"""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h264_decoder")
STREAMS = ["i_cabac", "i_high8", "i_slices", "i_crop", "i_cqm", "pcm_cabac", "c_8x8dct", "c_odd", "c_parts",
           "c_ref4", "c_slices", "c_weight", "ci_pc", "b_sp_cabac", "b_tm_cabac", "b_pyr", "b_wimp", "ci_b"]


@pytest.mark.parametrize("name", STREAMS)
def test_fixture_decodes_to_ffmpeg_pictures(oracle, name):
    from oracle import pydecoder

    frames = pydecoder.decode_mp4(os.path.join(HERE, f"h264_{name}.mp4"))
    want = np.fromfile(os.path.join(HERE, f"h264_{name}.ffmpeg.yuv"), np.uint8)
    h, w = frames[0].y.shape
    size = w * h + 2 * ((w + 1) // 2) * ((h + 1) // 2)
    assert len(want) == size * len(frames), f"{name}: {len(frames)} pictures, FFmpeg has {len(want) / size}"
    for k, f in enumerate(frames):
        got = np.concatenate([f.y.ravel(), f.cb.ravel(), f.cr.ravel()])
        assert np.array_equal(got, want[k * size:(k + 1) * size]), f"{name}: picture {k} differs from FFmpeg's"


def test_fixtures_exercise_every_prediction_form(oracle):
    """Across the fixtures, every Intra4x4 and Intra8x8 mode (the 8x8 reference filter with them), every
    Intra16x16 and chroma mode, every P partitioning and sub-partitioning, P_Skip, I_PCM, the 8x8
    transform in inter macroblocks and reference indexes above 0 occur: the oracle's primitives for
    each are pinned by FFmpeg's pictures."""
    from oracle import pydecoder

    tot = None
    for name in STREAMS:
        _, st = pydecoder.decode_mp4(os.path.join(HERE, f"h264_{name}.mp4"), with_stats=True)
        tot = st if tot is None else {k: tot[k] + st[k] for k in st}
    for k in ("i4_modes", "i8_modes", "i16_modes", "chroma_modes", "partitions", "sub_partitions"):
        assert np.all(np.asarray(tot[k]) > 0), (k, tot[k])
    for k in ("skip", "pcm", "inter_t8", "ref_gt0", "b_skip", "b_direct16", "b8x8", "b_temporal", "b_implicit"):
        assert tot[k] > 0, k
    # every B 16x16 direction, 16x8 and 8x16, and the 8x8 sub-types the encoder writes (direct, list 0,
    # list 1, both)
    assert np.all(np.asarray(tot["b16"]) > 0) and np.all(np.asarray(tot["b_partitions"]) > 0), (tot["b16"], tot["b_partitions"])
    assert np.all(np.asarray(tot["b_sub"][:4]) > 0), tot["b_sub"]
