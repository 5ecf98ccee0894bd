"""MI355X (gfx950) H.264 per-macroblock analysis stage for nervosys/IronAV's software encoder.

The compute runs in hand-written HIP kernels behind the C ABI of include/ironav_h264a.h
(libironav_h264a.so, built in-tree by ironav_amd.build); this package is the host side.
"""
from . import abi, clipgen  # noqa: F401
from .h264a import Analyser, Analysis, AnalysisError, Picture, library, slice_rows  # noqa: F401
