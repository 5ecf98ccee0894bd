"""Builds the gfx950 analysis library (ironav_amd/libironav_h264a.so) in-tree with hipcc.

Every HIP source is compiled for gfx950 only, with packed FP32 instructions disabled and without
floating-point contraction (the rate-distortion decisions are IEEE double arithmetic that must not
be fused), then linked into one shared library that exports the C ABI of include/ironav_h264a.h.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libironav_h264a.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.hip", "batch.hip", "halfpel.hip", "deblock.hip", "fullsearch.hip", "analyse.hip", "proposal.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-parameter",
    "-Wno-unused-command-line-argument",
]


def _compile(src, obj):
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{p.stderr}")
    return p.stderr


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, s.replace(".hip", ".o")) for s in SOURCES]
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        logs = list(ex.map(_compile, SOURCES, objs))
    if verbose:
        for s, log in zip(SOURCES, logs):
            if log.strip():
                print(f"[{s}]\n{log}", file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, OUT)
    return OUT


WRITER_OUT = os.path.join(HERE, "libironav_h264w.so")
WRITER_SOURCES = ["gop.c", "writer.c"]


def build_writer():
    """The host side around the analysis (include/ironav_h264w.h): picture-type decisions and the
    bitstream back end, plain C for the CPU, ironav_amd/libironav_h264w.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = os.environ.get("CC", "gcc")
    objs = []
    for src in WRITER_SOURCES:
        obj = os.path.join(objdir, src.replace(".c", ".w.o"))
        p = subprocess.run([cc, "-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-pthread",
                            "-c", os.path.join(CSRC, src), "-o", obj], capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"{cc} failed on {src}:\n{p.stderr}")
        objs.append(obj)
    tmp = WRITER_OUT + ".tmp"
    p = subprocess.run([cc, "-shared", "-pthread", "-o", tmp, *objs], capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, WRITER_OUT)
    return WRITER_OUT


TRAIN_OUT = os.path.join(HERE, "libironav_train.so")


def build_train():
    """The FP32 training kernels of the proposal CNN (offline tooling), ironav_amd/libironav_train.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    obj = os.path.join(objdir, "train.o")
    _compile("train.hip", obj)
    tmp = TRAIN_OUT + ".tmp"
    p = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, obj], capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, TRAIN_OUT)
    return TRAIN_OUT


PROBE_OUT = os.path.join(HERE, "libironav_probe.so")


def build_probe():
    """The VALU issue-rate measurement behind the SAD rooflines (`bench.py --config issue-rates`),
    ironav_amd/libironav_probe.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    obj = os.path.join(objdir, "issue_rates.o")
    _compile("issue_rates.hip", obj)
    tmp = PROBE_OUT + ".tmp"
    p = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, obj], capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, PROBE_OUT)
    return PROBE_OUT


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_train())
    print(build_writer())
    print(build_probe())
