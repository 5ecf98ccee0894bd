"""Throughput of the H.264 motion-analysis stage on one MI355X.

Workload (default "fullsearch-1080p"): one 1920x1080 P picture of the seeded synthetic clip searched
exhaustively (+-32 integer samples, all 41 partitions from 16x16 to 4x4) against 3 reference
pictures - the plain GPU full search the learned proposal is measured against. A step is one
picture (8,160 macroblocks x 3 references), inputs already resident in HBM. Metric: macroblocks per
second.

Prints one JSON line. The roofline object is that of the search kernel, timed with HIP events on the
stream it is launched on; its work is the algorithmic absolute-difference count of the clipped
windows. cpu_baseline is the oracle's scalar exhaustive search of the same picture against one
reference on one host core, scaled to three references.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_CU, MI355X_CLK_GHZ = 256, 2.4
# v_sad_u8 / v_sad_hi_u8: 4 absolute differences per lane per instruction, 64 lanes per CU per clock
VALU_ABSDIFF_PEAK_T = MI355X_CU * 64 * 4 * MI355X_CLK_GHZ / 1000.0   # 157.3 T abs-diff/s


def window_positions(mb_w, mb_h, rng):
    pw, ph = mb_w * 16, mb_h * 16
    tot = 0
    for mby in range(mb_h):
        ny = min(rng, ph - mby * 16 + 8) - max(-rng, -(mby * 16 + 24)) + 1
        for mbx in range(mb_w):
            nx = min(rng, pw - mbx * 16 + 8) - max(-rng, -(mbx * 16 + 24)) + 1
            tot += nx * ny
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--range", type=int, default=32)
    ap.add_argument("--refs", type=int, default=3)
    ap.add_argument("--variant", type=int, default=-1, help="full-search tile shape (-1: default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (streams or slice bands would "
              "shard across GPUs with no exchange; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ironav_amd import abi, clipgen
    from ironav_amd.h264a import Analyser, Picture, fullsearch_variants

    width, height, nref = 1920, 1080, args.refs
    mbs_ = ((width + 15) // 16) * ((height + 15) // 16)
    spec = clipgen.workload_1080p()
    spec.frames = nref + 1
    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]
    torch.cuda.init()
    stream = torch.cuda.Stream()  # default priority; the library enqueues on it, the events time it
    an = Analyser(width, height, nrefs=nref, search_mode=abi.SEARCH_FULL, search_range=args.range)
    an.set_stream(stream.cuda_stream)
    an.set_fullsearch_variant(args.variant)
    for r in range(nref):
        an.set_reference(r, pics[nref - 1 - r])   # list 0: most recent first
    an.upload(pics[nref])
    an.sync()
    refs = list(range(nref))
    for _ in range(args.warmup):
        an.motion_search(refs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        an.motion_search(refs)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_kernel = ev0.elapsed_time(ev1) / args.steps
    ms_step = max(ms_kernel, wall * 1000.0 / args.steps)
    mb_w, mb_h = an.mb_w, an.mb_h
    mbs = mb_w * mb_h
    value = mbs / (ms_step / 1000.0)
    absdiff = window_positions(mb_w, mb_h, args.range) * 256 * nref
    achieved = absdiff / (ms_kernel / 1000.0) / 1e12
    got = an.read_motion(nref)
    an.close()

    # the exact analysis (code_mb over every slice) of the same P picture, timed the same way
    analysis = {}
    for slices in (8, 68):
        with Analyser(width, height, nrefs=nref, slices=slices, search_mode=abi.SEARCH_FULL,
                      search_range=args.range) as az:
            az.set_stream(stream.cuda_stream)
            refres = az.analyse(pics[0], 25, abi.PIC_IDR)
            for r in range(nref):
                az.set_reference(r, pics[nref - 1 - r] if r else refres.recon, None if r else refres.mbs)
            az.upload(pics[nref])
            az.motion_search(refs)
            az.sync()
            reps = 3
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                az.motion_search(refs)
                az.analyse_uploaded(28, abi.PIC_P, refs)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            analysis[f"slices{slices}"] = {"ms_per_picture": round(ms, 3), "MB_per_s": round(mbs_ / (ms / 1000.0), 1)}

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle

        pyoracle.build()
        t = time.perf_counter()
        pyoracle.fullsearch(pics[nref].y, [pics[nref - 1].y], mb_w, mb_h, args.range)
        dt = time.perf_counter() - t
        cpu = {"value": round(mbs / (dt * nref), 1), "unit": "MB/s", "cores": 1, "kind": "port",
               "sample": f"oracle scalar exhaustive search of one 1080p picture against 1 reference "
                         f"(+-{args.range}, 41 partitions), {dt:.1f} s, scaled to {nref} references"}

    out = {
        "metric": "MB/s motion analysis @1080p: learned proposal vs full search at <=1% cost gap",
        "value": round(value, 1),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generator clip, 1920x1080 4:2:0, inputs resident in HBM)",
        "config": {"workload": "plain GPU exhaustive integer search, 1920x1080 P picture, "
                               f"+-{args.range}, 41 partitions, {nref} references (the full-search baseline; "
                               "the learned proposal is not built yet)",
                   "macroblocks_per_step": mbs, "references": nref, "search_range": args.range,
                   "kernel_variant": fullsearch_variants()[args.variant] if args.variant >= 0 else "default"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": round(VALU_ABSDIFF_PEAK_T, 1),
                     "unit": "Tabsdiff/s", "frac": round(achieved / VALU_ABSDIFF_PEAK_T, 4), "traffic": None,
                     "kernel": "fullsearch_kernel", "ms_per_launch": round(ms_kernel, 4),
                     "work_per_launch": absdiff},
        "cpu_baseline": cpu,
        "analysis": {"what": "exhaustive search + exact code_mb analysis of the same P picture (QP 28, "
                             "3 references, CABAC, 8x8 transform), per slice count", **analysis},
        "check": {"sad16_mean": float(np.mean(got["sad"][:, :, 0]))},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
