"""Copies the reference decoder's fixtures that the oracle's decoder is checked against.

The files are libx264 streams (MP4) with the pictures FFmpeg decodes them to (raw yuv420p), held by
nervosys/IronAV's engine tests (crates/ironav-engine/tests/data/, harness
crates/ironav-engine/tests/e2e.rs:1991-2036). The subset copied is the progressive 8-bit 4:2:0 CABAC
I and P streams, which is what the oracle's decoder decodes.

Usage: python tests/golden/h264_decoder/collect.py [reference root]
"""
import os
import shutil
import sys

NAMES = ["i_cabac", "i_high8", "i_slices", "i_crop", "i_cqm", "pcm_cabac", "c_8x8dct", "c_odd", "c_parts", "c_ref4",
         "c_slices", "c_weight", "ci_pc", "b_sp_cabac", "b_tm_cabac", "b_pyr", "b_wimp", "ci_b"]


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(root, "crates", "ironav-engine", "tests", "data")
    here = os.path.dirname(os.path.abspath(__file__))
    for n in NAMES:
        for ext in (".mp4", ".ffmpeg.yuv"):
            shutil.copyfile(os.path.join(src, f"h264_{n}{ext}"), os.path.join(here, f"h264_{n}{ext}"))


if __name__ == "__main__":
    main()
