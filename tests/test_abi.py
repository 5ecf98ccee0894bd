"""The C-ABI library loads on a host without a GPU and exports every entry point include/*.h declares."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from ironav_amd import abi, clipgen, h264a

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(ia_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    lib = h264a.library()
    names = declared_functions()
    assert "ia_create" in names and "ia_analyse_picture" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.PROTOTYPES), "abi.py must bind exactly the declared entry points"


def test_variants_listed_without_a_device():
    names = h264a.fullsearch_variants()
    assert len(names) >= 1 and all(names)


def test_create_rejects_bad_configs_before_touching_a_device():
    lib = h264a.library()
    h = C.c_void_p()
    bad = [
        abi.Config(15, 16, 8, 1, 1, 0, 0, 1, 1, 0, 16, 1.1),    # odd width
        abi.Config(64, 64, 10, 1, 1, 0, 0, 1, 1, 0, 16, 1.1),   # 10-bit
        abi.Config(64, 64, 8, 1, 1, 0, 0, 0, 1, 0, 16, 1.1),    # nrefs 0
        abi.Config(64, 64, 8, 1, 1, 0, 0, 1, 5, 0, 16, 1.1),    # more slices than MB rows
        abi.Config(64, 64, 8, 1, 1, 0, 0, 1, 1, 0, 65, 1.1),    # range
    ]
    for cfg in bad:
        rc = lib.ia_create(C.byref(h), C.byref(cfg))
        assert rc in (abi.IA_EINVAL, abi.IA_EUNSUPPORTED), rc
        assert lib.ia_last_error(None)


def test_clip_generator_is_deterministic_and_pinned():
    spec = clipgen.ClipSpec(64, 48, 3, 1234)
    a = [np.concatenate([p.ravel() for p in f]) for f in clipgen.make_clip(spec)]
    b = [np.concatenate([p.ravel() for p in f]) for f in clipgen.make_clip(spec)]
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    # motion: consecutive frames differ, and the picture is not flat
    assert not np.array_equal(a[0], a[1]) and a[0][: 64 * 48].std() > 10
    y, cb, cr = clipgen.pad_planes(*next(clipgen.make_clip(clipgen.ClipSpec(40, 26, 1, 5))))
    assert y.shape == (32, 48) and cb.shape == (16, 24)
    assert np.array_equal(y[26:, :40], np.repeat(y[25:26, :40], 6, axis=0))


def test_device_clip_generator_matches_host_without_noise():
    """make_clip_device (torch, here on the CPU) draws the same scene and motion as make_clip: with
    no noise and no cut the padded planes are identical."""
    torch = pytest.importorskip("torch")
    for spec in (clipgen.ClipSpec(96, 64, 3, 11, noise_sigma=0.0),
                 clipgen.ClipSpec(90, 50, 3, 12, noise_sigma=0.0, pan=3, max_motion=20)):
        host = [clipgen.pad_planes(*f) for f in clipgen.make_clip(spec)]
        dev = list(clipgen.make_clip_device(spec, device="cpu"))
        for fh, fd in zip(host, dev):
            for a, b in zip(fh, fd):
                assert isinstance(b, torch.Tensor) and np.array_equal(a, b.numpy())


def test_bd_rate_closed_forms():
    """Bjontegaard delta rate: equal curves give 0; a curve needing 10% more bits at every PSNR gives
    +10% (the cubic fit of log rate is exact for a constant offset); 10% fewer gives -10%."""
    from ironav_amd.bdrate import bd_rate

    psnr = [33.0, 35.5, 38.2, 41.0]
    rate = [1.0e5, 2.1e5, 4.6e5, 9.9e5]
    for method in ("pchip", "cubic"):
        assert abs(bd_rate(rate, psnr, rate, psnr, method)) < 1e-9
        assert abs(bd_rate(rate, psnr, [r * 1.1 for r in rate], psnr, method) - 10.0) < 1e-6
        assert abs(bd_rate(rate, psnr, [r * 0.9 for r in rate], psnr, method) + 10.0) < 1e-6
        # a curve shifted by +0.5 dB at the same rates needs fewer bits at equal quality
        assert bd_rate(rate, psnr, rate, [p + 0.5 for p in psnr], method) < 0


def test_realtime_settings_follow_the_encoders_live_mode():
    """h264a.realtime mirrors enc/mod.rs:249-254: the realtime decisions, one reference, a slice per
    thread up to one per two macroblock rows (at least one)."""
    assert h264a.realtime(1080, 64) == {"fast": True, "nrefs": 1, "slices": 34}
    assert h264a.realtime(1080, 16)["slices"] == 16
    assert h264a.realtime(16, 8)["slices"] == 1 and h264a.realtime(48, 1)["slices"] == 1
