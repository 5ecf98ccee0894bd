"""Builds the half-sample planes of one 1920x1080 reference picture repeatedly, for trace runs.

Each `ia_set_reference` uploads the luma and launches `halfpel_kernel` once. Beside it, two torch
kernels that move the same bytes give the floors at this size: a device copy of 5.82 MB (11.6 MB of
traffic, the kernel's algorithmic bytes) and a fill of the four planes (9.55 MB written).

    rocprofv3 --kernel-trace --stats -- python3 ironav_amd/probes/halfpel_loop.py --launches 100
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--launches", type=int, default=100)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()

    import torch
    from ironav_amd import clipgen
    from ironav_amd.h264a import Analyser, Picture

    w, h = args.width, args.height
    pic = Picture(*clipgen.pad_planes(*next(iter(clipgen.make_clip(clipgen.ClipSpec(w, h, 1, 21))))))
    with Analyser(w, h, nrefs=1) as an:
        for _ in range(args.launches):
            an.set_reference(0, pic)
    pw, ph = pic.y.shape[1], pic.y.shape[0]
    planes = 4 * (pw + 96) * (ph + 96)
    src = torch.empty((pw * ph + planes) // 2, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    fill = torch.empty(planes, dtype=torch.uint8, device="cuda")
    for _ in range(args.launches):
        dst.copy_(src)
        fill.fill_(7)
        torch.cuda.synchronize()
    print(f"{args.launches} builds of {w}x{h}; copy {src.numel()} B; fill {planes} B", flush=True)


if __name__ == "__main__":
    main()
