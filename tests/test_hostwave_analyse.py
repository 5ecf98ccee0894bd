"""The analysis kernel built for the host under AddressSanitizer and UBSan (bounds included), against
the oracle.

The kernel source (ironav_amd/csrc/analyse.hip) is compiled with g++ and tests/hostwave/hostwave.h,
which runs each 64-lane workgroup as 64 fibers on one host thread with the cross-lane operations as
barriers. Every index the kernel computes into its private arrays, its LDS fields and the global
buffers is checked: a sanitizer report or a diverged wave fails the run. The outputs must also be
bit-exact with the oracle, so the emulation runs the same code paths the GPU runs.

The cases include the IDR picture at QP 10 (96x64, 2 slices, CABAC, 8x8 transform) whose GPU run
once ended in a memory-aperture violation (see the design notes), IPPP chains with 1-3 references,
the exhaustive search's candidates, sub-8x8 partitions, CAVLC decisions and the realtime decisions.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ironav_amd import abi, clipgen
from ironav_amd.h264a import Picture

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hostwave", "analyse_host.cpp")
BUILD = os.path.join(HERE, "hostwave", "_build")
EXE = os.path.join(BUILD, "analyse_host")
CSRC = os.path.join(os.path.dirname(HERE), "ironav_amd", "csrc")


def build():
    deps = [SRC, os.path.join(HERE, "hostwave", "hostwave.h")] + [
        os.path.join(CSRC, f) for f in ("analyse.hip", "analyse.h", "common.h", "h264_tables.h")]
    if os.path.exists(EXE) and os.path.getmtime(EXE) >= max(os.path.getmtime(d) for d in deps):
        return EXE
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined,bounds-strict", "-fno-sanitize-recover=all",
           "-I" + os.path.join(HERE, "hostwave"), SRC, "-o", EXE + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(EXE + ".tmp", EXE)
    return EXE


@pytest.fixture(scope="module")
def host_exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    return build()


def frames(w, h, n, seed, **kw):
    spec = clipgen.ClipSpec(w, h, n, seed, **kw)
    return [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]


def run_host(exe, tmp_path, cfg, pic, qp, kind, refs=(), fs=None, halfpel=None, dist=(0, 0)):
    """refs: (Picture, records) in list 0 order; for a B picture (kind abi.PIC_B) the list 0 and list
    1 pictures, dist their distances (tb, td)."""
    mb_w, mb_h = (cfg.width + 15) // 16, (cfg.height + 15) // 16
    nmb = mb_w * mb_h
    nref = len(refs)
    col = refs[1] if kind == abi.PIC_B else refs[0] if refs else None
    mbs0 = col[1] if col is not None and col[1] is not None else None
    head = np.array([cfg.width, cfg.height, cfg.cabac, cfg.transform8, cfg.sub8x8, cfg.fast, cfg.slices,
                     cfg.search_mode, qp, kind, nref, mbs0 is not None, fs is not None, dist[0], dist[1]],
                    dtype=np.int32)
    parts = [head.tobytes(), pic.y.tobytes(), pic.cb.tobytes(), pic.cr.tobytes()]
    for r, _ in refs:
        parts += [halfpel(r.y).tobytes(), r.cb.tobytes(), r.cr.tobytes()]
    if mbs0 is not None:
        parts.append(np.ascontiguousarray(mbs0, dtype=abi.MBINFO_DTYPE).tobytes())
    if fs is not None:
        parts.append(np.ascontiguousarray(fs).tobytes())
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    inp.write_bytes(b"".join(parts))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, f"host build of the kernel failed (rc {p.returncode}):\n{p.stderr[-4000:]}"
    buf = out.read_bytes()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
        off += a.nbytes
        return a

    codes = take(abi.MBCODE_DTYPE, nmb)
    stats = take(abi.MBSTATS_DTYPE, nmb)
    ry = take(np.uint8, pic.y.size).reshape(pic.y.shape)
    rcb = take(np.uint8, pic.cb.size).reshape(pic.cb.shape)
    rcr = take(np.uint8, pic.cr.size).reshape(pic.cr.shape)
    mbs = take(abi.MBINFO_DTYPE, nmb)
    assert off == len(buf)
    return codes, stats, (ry, rcb, rcr), mbs


def compare(tag, got, want):
    gc, gs, grec, gm = got
    wc, ws, wrec, wm = want
    for f in abi.MBCODE_DTYPE.names:
        assert np.array_equal(gc[f], wc[f]), f"{tag}: code.{f} differs"
    for f in abi.MBSTATS_DTYPE.names:
        assert np.array_equal(gs[f], ws[f]), f"{tag}: stats.{f} differs"
    for k in range(3):
        assert np.array_equal(grec[k], wrec[k]), f"{tag}: reconstruction plane {k} differs"
    for f in abi.MBINFO_DTYPE.names:
        assert np.array_equal(gm[f], wm[f]), f"{tag}: mbinfo.{f} differs"


def chain(exe, tmp_path, oracle, w, h, nframes, qp, *, nrefs=1, slices=1, cabac=True, transform8=True,
          sub8x8=False, search_mode=abi.SEARCH_DIAMOND, search_range=16, fast=False, seed=1, **clip):
    pics = frames(w, h, nframes, seed, **clip)
    cfg = abi.Config(width=w, height=h, bit_depth=8, cabac=int(cabac), transform8=int(transform8),
                     sub8x8=int(sub8x8), fast=int(fast), nrefs=nrefs, slices=slices, search_mode=search_mode,
                     search_range=search_range, fallback_ratio=1.1)

    def halfpel(y):
        return oracle.halfpel_build(y.astype(np.uint16), 8).astype(np.uint8)

    refs = []
    mb_w, mb_h = (w + 15) // 16, (h + 15) // 16
    for t, pic in enumerate(pics):
        kind = abi.PIC_IDR if t == 0 else abi.PIC_P
        q = qp - 3 if kind == abi.PIC_IDR else qp
        nref = min(len(refs), nrefs) if kind == abi.PIC_P else 0
        fs = None
        if kind == abi.PIC_P and search_mode == abi.SEARCH_FULL:
            fs = oracle.fullsearch(pic.y, [r[0].y for r in refs[:nref]], mb_w, mb_h, search_range)
        got = run_host(exe, tmp_path, cfg, pic, q, kind, refs[:nref], fs, halfpel)
        want = oracle.analyse_picture(cfg, pic, q, kind, refs[:nref], fs)
        compare(f"frame {t} qp {q}", got, want)
        refs.insert(0, (Picture(*got[2]), got[3]))


@pytest.mark.parametrize("qp", [10, 28, 45])
def test_idr_picture_host(host_exe, tmp_path, oracle, qp):
    """The IDR case of the GPU tests at QP 10/28/45: 96x64, 2 slices, CABAC, 8x8 transform."""
    pic = frames(96, 64, 1, 5)[0]
    cfg = abi.Config(width=96, height=64, bit_depth=8, cabac=1, transform8=1, sub8x8=0, fast=0, nrefs=1,
                     slices=2, search_mode=abi.SEARCH_DIAMOND, search_range=16, fallback_ratio=1.1)
    got = run_host(host_exe, tmp_path, cfg, pic, qp, abi.PIC_IDR)
    compare(f"idr qp {qp}", got, oracle.analyse_picture(cfg, pic, qp, abi.PIC_IDR))


def test_ippp_chain_host(host_exe, tmp_path, oracle):
    chain(host_exe, tmp_path, oracle, 64, 48, 3, 22, nrefs=2, slices=2, seed=22)


def test_full_search_and_sub8x8_host(host_exe, tmp_path, oracle):
    chain(host_exe, tmp_path, oracle, 48, 48, 3, 18, nrefs=2, slices=1, sub8x8=True,
          search_mode=abi.SEARCH_FULL, search_range=12, seed=4, max_motion=8)


def test_cavlc_and_realtime_host(host_exe, tmp_path, oracle):
    chain(host_exe, tmp_path, oracle, 48, 32, 3, 26, nrefs=2, slices=2, cabac=False, transform8=False, seed=6)
    chain(host_exe, tmp_path, oracle, 48, 32, 3, 26, nrefs=1, slices=2, fast=True, seed=77, max_motion=6)


def b_chain(exe, tmp_path, oracle, w, h, nframes, qp, *, nrefs=2, slices=1, transform8=True, seed=1, **clip):
    """An IBBP chain (two B pictures, coding order of the reference's encoder); the P and B pictures
    analysed by the host build against the oracle, each reference the host build's reconstruction."""
    from ironav_amd import h264w

    pics = frames(w, h, nframes, seed, **clip)
    cfg = abi.Config(width=w, height=h, bit_depth=8, cabac=1, transform8=int(transform8), sub8x8=0, fast=0,
                     nrefs=nrefs, slices=slices, search_mode=abi.SEARCH_DIAMOND, search_range=16, fallback_ratio=1.1)

    def halfpel(y):
        return oracle.halfpel_build(y.astype(np.uint16), 8).astype(np.uint8)

    refs = []  # (display, Picture, records) in coding order
    nb = 0
    for d, kind in h264w.coding_order([t == 0 for t in range(nframes)], 2):
        pic = pics[d]
        if kind == abi.PIC_IDR:
            q, rl, dist = qp - 3, [], (0, 0)
        elif kind == abi.PIC_P:
            q, rl, dist = qp, [(r[1], r[2]) for r in refs[::-1][:nrefs]], (0, 0)
        else:
            l0 = max((r for r in refs if r[0] < d), key=lambda r: r[0])
            l1 = min((r for r in refs if r[0] > d), key=lambda r: r[0])
            q, rl, dist = min(qp + 5, 51), [(l0[1], l0[2]), (l1[1], l1[2])], (2 * (d - l0[0]), 2 * (l1[0] - l0[0]))
        got = run_host(exe, tmp_path, cfg, pic, q, kind, rl, None, halfpel, dist)
        if kind == abi.PIC_B:
            want = oracle.analyse_picture_b(cfg, pic, q, rl[0], rl[1], *dist)
            nb += 1
        else:
            want = oracle.analyse_picture(cfg, pic, q, kind, rl)
        compare(f"picture {d} kind {kind} qp {q}", got, want)
        if kind != abi.PIC_B:
            if len(refs) == max(nrefs, 2):
                refs.pop(0)
            refs.append((d, Picture(*got[2]), got[3]))
    assert nb > 0


def test_b_pictures_host(host_exe, tmp_path, oracle):
    b_chain(host_exe, tmp_path, oracle, 64, 48, 5, 26, nrefs=2, slices=2, seed=31, max_motion=6)
    b_chain(host_exe, tmp_path, oracle, 48, 32, 4, 34, nrefs=1, slices=1, transform8=False, seed=32)
