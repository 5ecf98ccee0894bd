"""Seeded synthetic 8-bit 4:2:0 clips for the analysis stage.

Luma is a sum of random-phase 2-D sinusoids (periods 8..128 samples) plus textured rectangles that
each move by a per-object integer vector every frame (plus an optional global pan), plus Gaussian
noise, clipped to 0..255. Chroma is the 2x2 average of luma with an offset. After the sine table
(rounded to integers), everything is integer arithmetic on numpy's PCG64 stream, so a (size, seed,
frame) triple gives the same bytes on every host with this numpy.

`make_clip` returns unpadded planes; `pad_planes` repeats the right column and bottom row up to
whole macroblocks as the encoder's `source()` does (crates/ironav-codec/src/h264/enc/mod.rs:481-529).
"""
from dataclasses import dataclass, field

import numpy as np

_LUT_BITS = 10
_LUT = np.round(np.sin(2.0 * np.pi * np.arange(1 << _LUT_BITS) / (1 << _LUT_BITS)) * 1024.0).astype(np.int32)


@dataclass
class ClipSpec:
    width: int
    height: int
    frames: int
    seed: int
    n_waves: int = 6
    n_rects: int = 4
    max_motion: int = 12          # per-object integer motion, uniform in +-max_motion per frame
    pan: int = 0                  # global pan, uniform in +-pan per frame (fixed per clip)
    noise_sigma: float = 2.0
    cuts: tuple = field(default_factory=tuple)  # frames at which the scene is redrawn


def _scene(rng, spec):
    waves = []
    for _ in range(spec.n_waves):
        px, py = rng.integers(8, 129, size=2)
        sx, sy = rng.choice([-1, 1], size=2)
        # frequency in 1/65536 cycles per sample; one axis may be nearly flat
        fx = int(sx * round(65536 / px)) if rng.random() < 0.8 else 0
        fy = int(sy * round(65536 / py)) if (fx == 0 or rng.random() < 0.8) else 0
        waves.append((fx, fy, int(rng.integers(0, 65536)), int(rng.integers(6, 19))))
    rects = []
    for _ in range(spec.n_rects):
        rw = int(rng.integers(max(16, spec.width // 16), max(17, spec.width // 4)))
        rh = int(rng.integers(max(16, spec.height // 16), max(17, spec.height // 4)))
        x0 = int(rng.integers(0, spec.width))
        y0 = int(rng.integers(0, spec.height))
        vx, vy = (int(v) for v in rng.integers(-spec.max_motion, spec.max_motion + 1, size=2))
        tfx, tfy = (int(v) for v in rng.integers(1200, 9000, size=2) * rng.choice([-1, 1], size=2))
        base = int(rng.integers(40, 216))
        amp = int(rng.integers(20, 60))
        rects.append((rw, rh, x0, y0, vx, vy, tfx, tfy, base, amp))
    pan = tuple(int(v) for v in rng.integers(-spec.pan, spec.pan + 1, size=2)) if spec.pan else (0, 0)
    return waves, rects, pan


def _luma(spec, scene, t, rng):
    waves, rects, (panx, pany) = scene
    W, H = spec.width, spec.height
    ys, xs = np.mgrid[0:H, 0:W].astype(np.int64)
    bx, by = xs + panx * t, ys + pany * t
    acc = np.full((H, W), 128 << _LUT_BITS, dtype=np.int64)
    for fx, fy, ph, amp in waves:
        idx = ((fx * bx + fy * by + ph) & 0xFFFF) >> (16 - _LUT_BITS)
        acc += amp * _LUT[idx]
    img = acc >> _LUT_BITS
    for rw, rh, x0, y0, vx, vy, tfx, tfy, base, amp in rects:
        u = (xs - (x0 + vx * t)) % W
        v = (ys - (y0 + vy * t)) % H
        inside = (u < rw) & (v < rh)
        phase = ((tfx * u + tfy * v) & 0xFFFF) >> (16 - _LUT_BITS)
        checker = (((u >> 3) ^ (v >> 3)) & 1) * 24 - 12
        tex = base + ((amp * _LUT[phase]) >> _LUT_BITS) + checker
        img = np.where(inside, tex, img)
    noise = np.rint(rng.normal(0.0, spec.noise_sigma, size=(H, W))).astype(np.int64)
    return np.clip(img + noise, 0, 255).astype(np.uint8)


def _chroma(y):
    h, w = y.shape
    q = y.astype(np.int32)
    avg = (q[0::2, 0::2] + q[1::2, 0::2] + q[0::2, 1::2] + q[1::2, 1::2] + 2) >> 2
    cb = np.clip((avg >> 1) + 72, 0, 255).astype(np.uint8)
    cr = np.clip(184 - (avg >> 1), 0, 255).astype(np.uint8)
    return cb, cr


def make_clip(spec: ClipSpec):
    """Yields (y, cb, cr) uint8 planes of size height x width and height/2 x width/2 per frame."""
    if spec.width % 2 or spec.height % 2:
        raise ValueError("width and height must be even")
    rng = np.random.Generator(np.random.PCG64(spec.seed))
    scene = _scene(rng, spec)
    t0 = 0
    for t in range(spec.frames):
        if t in spec.cuts:
            scene = _scene(rng, spec)
            t0 = t
        y = _luma(spec, scene, t - t0, rng)
        cb, cr = _chroma(y)
        yield y, cb, cr


def make_clip_device(spec: ClipSpec, device="cuda"):
    """The same kind of clips made on a device with torch, for workloads too large to generate on the
    host: the first scene and its motion are make_clip's (the same seeded draws); the noise comes
    from torch's generator on the device, and a scene after a cut from a generator seeded with
    (seed, cut frame), so frames equal make_clip's up to the noise until the first cut. Yields planes
    padded to whole macroblocks (as pad_planes) as contiguous uint8 tensors on the device."""
    import torch

    if spec.width % 2 or spec.height % 2:
        raise ValueError("width and height must be even")
    rng = np.random.Generator(np.random.PCG64(spec.seed))
    scene = _scene(rng, spec)
    gen = torch.Generator(device=device)
    gen.manual_seed(spec.seed)
    lut = torch.tensor(_LUT, dtype=torch.int64, device=device)
    W, H = spec.width, spec.height
    pw, ph = (W + 15) // 16 * 16, (H + 15) // 16 * 16
    ys = torch.arange(H, dtype=torch.int64, device=device).view(H, 1)
    xs = torch.arange(W, dtype=torch.int64, device=device).view(1, W)
    t0 = 0
    for t in range(spec.frames):
        if t in spec.cuts:
            scene = _scene(np.random.Generator(np.random.PCG64([spec.seed, t])), spec)
            t0 = t
        tt = t - t0
        waves, rects, (panx, pany) = scene
        bx, by = xs + panx * tt, ys + pany * tt
        acc = torch.full((H, W), 128 << _LUT_BITS, dtype=torch.int64, device=device)
        for fx, fy, phs, amp in waves:
            acc += amp * lut[((fx * bx + fy * by + phs) & 0xFFFF) >> (16 - _LUT_BITS)]
        img = acc >> _LUT_BITS
        for rw, rh, x0, y0, vx, vy, tfx, tfy, base, amp in rects:
            u = torch.remainder(xs - (x0 + vx * tt), W).expand(H, W)
            v = torch.remainder(ys - (y0 + vy * tt), H).expand(H, W)
            inside = (u < rw) & (v < rh)
            phase = ((tfx * u + tfy * v) & 0xFFFF) >> (16 - _LUT_BITS)
            checker = (((u >> 3) ^ (v >> 3)) & 1) * 24 - 12
            tex = base + ((amp * lut[phase]) >> _LUT_BITS) + checker
            img = torch.where(inside, tex, img)
        noise = torch.round(torch.randn((H, W), generator=gen, device=device) * spec.noise_sigma).to(torch.int64)
        y = torch.clamp(img + noise, 0, 255).to(torch.int32)
        avg = (y[0::2, 0::2] + y[1::2, 0::2] + y[0::2, 1::2] + y[1::2, 1::2] + 2) >> 2
        cb = torch.clamp((avg >> 1) + 72, 0, 255)
        cr = torch.clamp(184 - (avg >> 1), 0, 255)

        def pad(a, h2, w2):
            a = torch.cat([a, a[:, -1:].expand(a.shape[0], w2 - a.shape[1])], 1) if w2 > a.shape[1] else a
            a = torch.cat([a, a[-1:, :].expand(h2 - a.shape[0], a.shape[1])], 0) if h2 > a.shape[0] else a
            return a.to(torch.uint8).contiguous()

        yield pad(y, ph, pw), pad(cb, ph // 2, pw // 2), pad(cr, ph // 2, pw // 2)


def pad_planes(y, cb, cr):
    """Pads planes to whole macroblocks by repeating the last column and row."""
    h, w = y.shape
    pw, ph = (w + 15) // 16 * 16, (h + 15) // 16 * 16
    py = np.pad(y, ((0, ph - h), (0, pw - w)), mode="edge")
    pcb = np.pad(cb, ((0, ph // 2 - h // 2), (0, pw // 2 - w // 2)), mode="edge")
    pcr = np.pad(cr, ((0, ph // 2 - h // 2), (0, pw // 2 - w // 2)), mode="edge")
    return np.ascontiguousarray(py), np.ascontiguousarray(pcb), np.ascontiguousarray(pcr)


# The clips of the workloads (sizes, lengths, motion and noise as the workloads state them).
def workload_720p(seed=720):
    return ClipSpec(1280, 720, 30, seed, max_motion=12, noise_sigma=2.0)


def workload_1080p(seed=1080):
    return ClipSpec(1920, 1080, 60, seed, max_motion=12, noise_sigma=2.0)
