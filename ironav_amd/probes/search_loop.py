"""Launches the motion search of the 1080p workload picture repeatedly, for counter and trace runs.

The 1920x1080 generator clip: pictures 0-3 are coded IPPP by the analysis (so reference 0 carries the
encoder's vectors), picture 4 is searched against the 3 references, `--launches` times, by the
exhaustive search (`--mode full`, +-32) or the learned proposal with the fallback off
(`--mode learned`: the proposal kernel alone), or on at 1.1x (`--mode learned-fallback`: the
proposal, then the exhaustive search of its fallback list).

    rocprofv3 --pmc ... -- python3 ironav_amd/probes/search_loop.py --mode full --launches 5
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=["full", "learned", "learned-fallback"], default="full")
    ap.add_argument("--launches", type=int, default=5)
    args = ap.parse_args()

    import torch  # noqa: F401  (the HIP runtime torch loads is the one the library binds to)
    from ironav_amd import abi, clipgen, proposal
    from ironav_amd.h264a import Analyser, Picture

    spec = clipgen.workload_1080p()
    spec.frames = 5
    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]
    coder = Analyser(1920, 1080, nrefs=3, slices=68, search_mode=abi.SEARCH_FULL, search_range=32)
    refs = []
    for t in range(4):
        n = min(len(refs), 3)
        for s in range(n):
            coder.set_reference(s, refs[s][0], refs[s][1])
        res = coder.analyse(pics[t], 25 if t == 0 else 28, abi.PIC_IDR if t == 0 else abi.PIC_P, list(range(n)))
        refs.insert(0, (res.recon, res.mbs))
    coder.close()
    if args.mode == "full":
        an = Analyser(1920, 1080, nrefs=3, slices=68, search_mode=abi.SEARCH_FULL, search_range=32)
    else:
        an = Analyser(1920, 1080, nrefs=3, slices=68, search_mode=abi.SEARCH_LEARNED, search_range=32,
                      fallback_ratio=1.1 if args.mode == "learned-fallback" else float("inf"))
        an.load_proposal_weights(proposal.load_weights())
    for s in range(3):
        an.set_reference(s, refs[s][0], refs[s][1])
    an.upload(pics[4])
    for _ in range(args.launches):
        an.motion_search([0, 1, 2])
    an.sync()
    an.close()
    print(f"{args.launches} {args.mode} searches done", flush=True)


if __name__ == "__main__":
    main()
