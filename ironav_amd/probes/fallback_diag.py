"""Why the learned proposal falls back: its reach against its cost prediction, on generator content.

Codes pictures 0-3 of a 1920x1080 clip (objects +-64 px/picture, noise sigma --sigma) IPPP with the
exhaustive search, then searches picture 4 against 3 references three ways: the +-64 exhaustive
search, the learned proposal with its fallback at 1.1x, and the proposal alone (fallback off). Per
reference it prints: the fallback rate; the share of exhaustive winners on the proposal's grid;
the share of pairs where the proposal alone finds the exhaustive SAD; and, over those pairs, the
residual log2(1 + refined SAD) - predicted log2 SAD (the 1.1x test falls back above 0.1375).

    python3 ironav_amd/probes/fallback_diag.py --sigma 4
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sigma", type=float, default=4.0)
    ap.add_argument("--motion", type=int, default=64)
    args = ap.parse_args()

    import numpy as np
    import torch  # noqa: F401  (the HIP runtime torch loads is the one the library binds to)

    from ironav_amd import abi, clipgen, proposal
    from ironav_amd.h264a import Analyser, Picture

    w, h, rng = 1920, 1080, 64
    spec = clipgen.ClipSpec(w, h, 5, 8001, max_motion=args.motion, noise_sigma=args.sigma)
    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]
    refs = []
    with Analyser(w, h, nrefs=3, slices=68, search_mode=abi.SEARCH_FULL, search_range=rng) as coder:
        for t in range(4):
            n = min(len(refs), 3)
            for s in range(n):
                coder.set_reference(s, refs[s][0], refs[s][1])
            res = coder.analyse(pics[t], 25 if t == 0 else 28, abi.PIC_IDR if t == 0 else abi.PIC_P, list(range(n)))
            refs.insert(0, (res.recon, res.mbs))
    params = proposal.load_weights()
    out = {}
    for name, mode, ratio in (("full", abi.SEARCH_FULL, 1.1), ("learned", abi.SEARCH_LEARNED, 1.1),
                              ("alone", abi.SEARCH_LEARNED, float("inf"))):
        with Analyser(w, h, nrefs=3, search_mode=mode, search_range=rng, fallback_ratio=ratio) as an:
            if mode == abi.SEARCH_LEARNED:
                an.load_proposal_weights(params)
            for s in range(3):
                an.set_reference(s, refs[s][0], refs[s][1])
            an.upload(pics[4])
            an.motion_search([0, 1, 2])
            out[name] = an.read_motion(3)
            if name == "learned":
                out["fallbacks"] = an.fallback_count()
            if name == "alone":
                feat, ctr = an.proposal_features([0, 1, 2])
                cnn = an.proposal_forward(feat)
    F, N = out["full"], out["alone"]
    cells, _ = proposal.labels(F, ctr.reshape(-1, 3, 2), 3)
    nmb = F.shape[1]
    print(f"sigma {args.sigma}, objects +-{args.motion}: fallback {out['fallbacks'] / (3 * nmb):.3f} of the pairs")
    for r in range(3):
        best = np.argmax(cnn[:, :, r], axis=1)
        pred = cnn[np.arange(nmb), best, 3]
        refined = np.log2(1.0 + N["sad"][r, :, 0].astype(np.float64))
        found = N["sad"][r, :, 0] <= F["sad"][r, :, 0]
        resid = (refined - pred)[found]
        print(f"  reference {r}: winners on the grid {np.mean(cells[:, r] >= 0):.3f}; found {found.mean():.3f}; "
              f"residual over found pairs p50 {np.median(resid):+.3f} p90 {np.percentile(resid, 90):+.3f}", flush=True)


if __name__ == "__main__":
    main()
