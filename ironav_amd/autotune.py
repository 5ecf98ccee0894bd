"""The autotuner of the exhaustive search's tile shape: per picture size and search range, every
compiled shape (strip rows K, workgroup width, waves per SIMD; see fullsearch.hip) is timed on a
generated picture pair against one reference, and the fastest is kept. The window class (±32 or
±64 staged rows) follows from the range. The table, ironav_amd/autotune.json, maps "WxH:R" to the
chosen shape with every shape's time; `apply` sets a context's shape from it (the batched path
then groups streams by shape, one launch per shape).

    python -m ironav_amd.autotune          # on the GPU; writes the table ($IA_AUTOTUNE_OUT overrides)
"""
import json
import os

from . import abi, clipgen
from .h264a import Analyser, Picture, fullsearch_variants

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "autotune.json")
SIZES = [(640, 360), (854, 480), (1280, 720), (1920, 1080), (3840, 2160)]
RANGES = [16, 32, 64]


def key(width, height, search_range):
    return f"{width}x{height}:{search_range}"


def tune(sizes=SIZES, ranges=RANGES, reps=10, log=None):
    """Times every tile shape at every (size, range); returns the table."""
    import torch

    names = fullsearch_variants()
    table = {}
    for w, h in sizes:
        spec = clipgen.ClipSpec(w, h, 2, 7700 + w, max_motion=24, pan=4)
        frames = [Picture(*(p.cpu().numpy() for p in f)) for f in clipgen.make_clip_device(spec)]
        for r in ranges:
            with Analyser(w, h, nrefs=1, search_mode=abi.SEARCH_FULL, search_range=r) as an:
                stream = torch.cuda.Stream()
                an.set_stream(stream.cuda_stream)
                an.set_reference(0, frames[0])
                an.upload(frames[1])
                an.sync()
                times = {}
                for v, name in enumerate(names):
                    an.set_fullsearch_variant(v)
                    for _ in range(2):
                        an.motion_search([0])
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    for _ in range(reps):
                        an.motion_search([0])
                    e1.record(stream)
                    e1.synchronize()
                    times[name] = round(e0.elapsed_time(e1) / reps, 4)
                best = min(times, key=times.get)
                table[key(w, h, r)] = {"variant": best, "ms": times}
                if log:
                    log(f"{key(w, h, r)}: {best} {times[best]} ms")
    return table


def load(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def choose(width, height, search_range, table=None):
    """The tuned shape's index for this size and range: the entry of the same range whose size is
    nearest in macroblocks; None when the table has none for the range."""
    table = load() if table is None else table
    names = fullsearch_variants()
    best, dist = None, None
    mbs = ((width + 15) // 16) * ((height + 15) // 16)
    for k, e in table.items():
        size, r = k.split(":")
        if int(r) != search_range or e["variant"] not in names:
            continue
        tw, th = (int(v) for v in size.split("x"))
        d = abs(((tw + 15) // 16) * ((th + 15) // 16) - mbs)
        if dist is None or d < dist:
            best, dist = names.index(e["variant"]), d
    return best


def apply(an, table=None):
    """Sets an Analyser's exhaustive-search shape from the table (unchanged when it has no entry)."""
    v = choose(an.cfg.width, an.cfg.height, an.cfg.search_range, table)
    if v is not None:
        an.set_fullsearch_variant(v)
    return v


def main():
    import time

    t0 = time.time()
    table = tune(log=lambda m: print(f"[{time.time() - t0:6.1f}s] {m}", flush=True))
    out = os.environ.get("IA_AUTOTUNE_OUT", TABLE)
    with open(out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
