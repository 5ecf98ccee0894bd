"""Rate-distortion of the learned proposal against the exhaustive search: the Bjontegaard delta rate.

Each search mode codes the same generator clip IPPP at four QPs, closed loop: each chain's
references are its own reconstructions, loop-filtered on the device as the encoder keeps them
(`ia_promote_reference`, enc/mod.rs:653-667). The rate of a picture is the size of its access unit as
the host back end writes it (libironav_h264w: slice headers and CABAC slice data of the analysed
macroblock codes, enc/frame.rs:2020-2279); the model rate the analysis counted
(`ia_mbstats.bits`, `CabacCost`, enc/cabac.rs:150-191) is reported beside it. Distortion is the PSNR
of the loop-filtered picture over the picture's visible area, luma and chroma weighted 4:1:1: the
picture a decoder rebuilds from the written stream (the GPU tests decode such chains and compare them
with these references byte for byte).

BD-rate (Bjontegaard): log rate interpolated over PSNR per mode (piecewise cubic Hermite, as JVET's
common test conditions compute it; the single-cubic form of VCEG-M33 beside it), the mean difference
over the PSNR range both curves cover, as a percentage: negative = the learned proposal needs fewer
bits for the same quality.

    python -m ironav_amd.bdrate                  # on the GPU: 1080p (+-32) and 2160p-type (+-64) clips
"""
import json
import sys

import numpy as np

QPS = (22, 27, 32, 37)


def bd_rate(rate_a, psnr_a, rate_b, psnr_b, method="pchip"):
    """Bjontegaard delta rate of curve b against curve a, in percent: log rate interpolated over PSNR,
    its mean difference over the PSNR range both curves cover. "pchip": piecewise cubic Hermite
    interpolation (the form JVET's common test conditions use; stable when a curve bends sharply);
    "cubic": one cubic polynomial per curve (VCEG-M33)."""
    pa, pb = np.asarray(psnr_a, float), np.asarray(psnr_b, float)
    la, lb = np.log(np.asarray(rate_a, float)), np.log(np.asarray(rate_b, float))
    lo, hi = max(pa.min(), pb.min()), min(pa.max(), pb.max())
    if method == "cubic":
        ia, ib = np.polyint(np.polyfit(pa, la, 3)), np.polyint(np.polyfit(pb, lb, 3))
        da = np.polyval(ia, hi) - np.polyval(ia, lo)
        db = np.polyval(ib, hi) - np.polyval(ib, lo)
    else:
        from scipy.interpolate import PchipInterpolator

        oa, ob = np.argsort(pa), np.argsort(pb)
        da = PchipInterpolator(pa[oa], la[oa]).integrate(lo, hi)
        db = PchipInterpolator(pb[ob], lb[ob]).integrate(lo, hi)
    return float((np.exp((db - da) / (hi - lo)) - 1.0) * 100.0)


def psnr(src, rec, w, h):
    """PSNR over the visible area, luma and chroma weighted 4:1:1 (MSE averaged with those weights)."""
    def mse(a, b, ww, hh):
        d = a[:hh, :ww].astype(np.float64) - b[:hh, :ww].astype(np.float64)
        return float(np.mean(d * d))
    m = (4 * mse(src.y, rec.y, w, h) + mse(src.cb, rec.cb, w // 2, h // 2) + mse(src.cr, rec.cr, w // 2, h // 2)) / 6
    return 10.0 * np.log10(255.0 * 255.0 / max(m, 1e-10))


def code_chain(pics, w, h, qp, mode, search_range, params=None, nrefs=3, slices=None):
    """Codes the pictures IPPP (IDR at qp - 3, as the reference's driver) with one search mode; returns
    (written bits, model bits, mean PSNR) over the P pictures."""
    from . import abi, h264w
    from .h264a import Analyser

    slices = slices or (h + 15) // 16
    bits, model, ps = 0.0, 0.0, []
    with Analyser(w, h, nrefs=nrefs, slices=slices, search_mode=mode, search_range=search_range,
                  fallback_ratio=1.1) as an, h264w.Stream(w, h, nrefs=nrefs, slices=slices) as wr:
        if params is not None:
            an.load_proposal_weights(params)
        order = []  # reference slots, newest picture first
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            q = max(0, qp - 3) if t == 0 else qp
            res = an.analyse(pic, q, kind, order[:nrefs])
            au = wr.write(kind, q, res.codes, len(order[:nrefs]))
            slot = order.pop() if len(order) == nrefs else len(order)
            an.promote_reference(slot, deblock=True)
            order.insert(0, slot)
            if t > 0:
                bits += 8.0 * len(au)
                model += float(res.stats["bits"].sum()) / 256.0
                ps.append(psnr(pic, an.read_reference(slot)[0], w, h))
    return bits, model, float(np.mean(ps))


def run(name, spec, search_range, log=print):
    from . import abi, clipgen, proposal
    from .h264a import Picture

    pics = [Picture(*(p.cpu().numpy() for p in f)) for f in clipgen.make_clip_device(spec)]
    params = proposal.load_weights()
    curves, models = {}, {}
    for label, mode in (("exhaustive", abi.SEARCH_FULL), ("learned", abi.SEARCH_LEARNED)):
        pts, mods = [], []
        for qp in QPS:
            b, m, p = code_chain(pics, spec.width, spec.height, qp, mode, search_range,
                                 params if mode == abi.SEARCH_LEARNED else None)
            pts.append((b, p))
            mods.append(m)
            log(f"[bdrate] {name} {label} qp {qp}: {b / max(1, len(pics) - 1):.0f} bits/picture written "
                f"({m / max(1, len(pics) - 1):.0f} modelled), {p:.3f} dB")
        curves[label] = pts
        models[label] = mods
    ra, pa = zip(*curves["exhaustive"])
    rb, pb = zip(*curves["learned"])
    return {"clip": name, "qps": list(QPS), "pictures": len(pics), "search_range": search_range,
            "exhaustive": {"bits": [round(x, 1) for x in ra], "model_bits": [round(x, 1) for x in models["exhaustive"]],
                           "psnr": [round(x, 4) for x in pa]},
            "learned": {"bits": [round(x, 1) for x in rb], "model_bits": [round(x, 1) for x in models["learned"]],
                        "psnr": [round(x, 4) for x in pb]},
            "bd_rate_percent": round(bd_rate(ra, pa, rb, pb), 3),
            "bd_rate_percent_model_rate": round(bd_rate(models["exhaustive"], pa, models["learned"], pb), 3),
            "bd_rate_percent_cubic": round(bd_rate(ra, pa, rb, pb, "cubic"), 3)}


def main():
    from . import clipgen

    log = lambda m: print(m, file=sys.stderr, flush=True)
    s1 = clipgen.workload_1080p()
    s1.frames = 8
    print(json.dumps(run("1920x1080 workload clip (objects +-12, noise 2), 8 pictures", s1, 32, log)), flush=True)
    s2 = clipgen.ClipSpec(3840, 2160, 8, 8000, max_motion=64, noise_sigma=4.0)
    print(json.dumps(run("3840x2160 (objects +-64, noise 4), 8 pictures", s2, 64, log)), flush=True)


if __name__ == "__main__":
    main()
