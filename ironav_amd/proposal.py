"""The learned motion proposal: parameters, dataset, training and evaluation.

The CNN (4 x 3x3 convolutions, 32 channels, on a 9x9 grid of 4-sample displacement cells around
each reference's temporal predictor) runs in BF16 on MFMA inside the analysis library
(ia_motion_search with IA_SEARCH_LEARNED). Training runs the FP32 HIP kernels of
libironav_train.so; torch supplies only device memory and the stream.

Labels are the exhaustive search's own 16x16 winners (ia_motion_search with IA_SEARCH_FULL) on
generator clips: the cell whose +-2 refinement window holds the winner, and log2(1 + SAD) of it.

Parameters are one float32 vector of abi.CNN_PARAMS values: per layer, weights [cout][cin][3][3]
then biases [cout], (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 4).
"""
import ctypes as C
import os

import numpy as np

from . import abi, clipgen
from .h264a import Analyser, Picture

LAYERS = [(8, 32), (32, 32), (32, 32), (32, 4)]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposal.npz")
TRAIN_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libironav_train.so")


def param_slices():
    out, o = [], 0
    for cin, cout in LAYERS:
        w = (o, o + cout * cin * 9)
        o = w[1]
        b = (o, o + cout)
        o = b[1]
        out.append((w, b))
    assert o == abi.CNN_PARAMS
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(abi.CNN_PARAMS, dtype=np.float32)
    for (w, b), (cin, cout) in zip(param_slices(), LAYERS):
        p[w[0]:w[1]] = rng.normal(0.0, np.sqrt(2.0 / (cin * 9)), size=w[1] - w[0])
    return p


def load_weights(path=WEIGHTS):
    with np.load(path) as z:
        return np.asarray(z["params"], dtype=np.float32)


def forward_reference(params, feat):
    """FP32 forward of the CNN in numpy (for checks): feat [n][81][8] -> [n][81][4]."""
    x = feat.reshape(-1, 9, 9, 8).astype(np.float64)
    for li, ((w, b), (cin, cout)) in enumerate(zip(param_slices(), LAYERS)):
        W = params[w[0]:w[1]].reshape(cout, cin, 3, 3).astype(np.float64)
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        y = np.zeros(x.shape[:3] + (cout,))
        for ty in range(3):
            for tx in range(3):
                y += xp[:, ty:ty + 9, tx:tx + 9, :] @ W[:, :, ty, tx].T
        y += params[b[0]:b[1]]
        x = np.maximum(y, 0.0) if li < 3 else y
    return x.reshape(-1, 81, 4).astype(np.float32)


def labels(fs, centres, nref):
    """Target cell [mb][3] (-1 when the winner lies outside the grid) and log2(1 + SAD) [mb][3] from
    exhaustive-search results fs [nref][mb][41] and grid centres [mb][3][2]."""
    n = fs.shape[1]
    cell = np.full((n, 3), -1, dtype=np.int32)
    cost = np.zeros((n, 3), dtype=np.float32)
    for r in range(nref):
        mv = fs[r]["mv"][:, 0, :].astype(np.int64)  # quarter samples, integer vectors
        px = mv // 4
        c = np.rint((px - centres[:, r, :]) / 4.0).astype(np.int64) + 4
        ok = np.all((c >= 0) & (c <= 8), axis=1)
        cell[:, r] = np.where(ok, c[:, 1] * 9 + c[:, 0], -1)
        cost[:, r] = np.log2(1.0 + fs[r]["sad"][:, 0].astype(np.float64))
    return cell, cost


def make_dataset(n_clips=200, width=320, height=192, frames=6, seed=20000, search_range=32, motion=(4, 25),
                 pan=(0, 5), noise=(1.0, 4.0), log=None):
    """Features, target cells and costs from generator clips: each clip coded IPPP (analysis on the
    GPU, so reference 0 carries the encoder's own vectors), each P picture searched exhaustively.
    Per clip: object motion bound, pan and noise sigma drawn from [motion), [pan) and [noise)."""
    feats, cells, costs = [], [], []
    with Analyser(width, height, nrefs=3, slices=max(1, (height + 15) // 16), search_mode=abi.SEARCH_FULL,
                  search_range=search_range) as an:
        for k in range(n_clips):
            rng = np.random.default_rng(seed + k)
            spec = clipgen.ClipSpec(width, height, frames, seed + k, max_motion=int(rng.integers(*motion)),
                                    pan=int(rng.integers(*pan)), noise_sigma=float(rng.uniform(*noise)))
            pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(spec)]
            refs = []
            for t, pic in enumerate(pics):
                if t == 0:
                    res = an.analyse(pic, 25, abi.PIC_IDR)
                else:
                    nref = min(3, len(refs))
                    for s in range(nref):
                        an.set_reference(s, refs[s][0], refs[s][1])
                    an.upload(pic)
                    feat, ctr = an.proposal_features(list(range(nref)))
                    an.motion_search(list(range(nref)))
                    fs = an.read_motion(nref)
                    cell, cost = labels(fs, ctr, nref)
                    feats.append(feat)
                    cells.append(cell)
                    costs.append(cost)
                    res = an.analyse(pic, int(rng.integers(22, 35)), abi.PIC_P, list(range(nref)))
                refs.insert(0, (res.recon, res.mbs))
            if log and k % 20 == 0:
                log(f"dataset: clip {k}/{n_clips}")
    return np.concatenate(feats), np.concatenate(cells), np.concatenate(costs)


class _TrainLib:
    def __init__(self):
        if not os.path.exists(TRAIN_LIB):
            raise RuntimeError(f"{TRAIN_LIB} is missing: build it with `python -m ironav_amd.build`")
        L = C.CDLL(TRAIN_LIB)
        vp, i = C.c_void_p, C.c_int
        for name, args in {
            "it_conv_fwd": [vp, vp, vp, vp, i, i, i, i, vp],
            "it_conv_bwd_data": [vp, vp, vp, i, i, i, vp],
            "it_conv_bwd_w": [vp, vp, vp, vp, i, i, i, vp],
            "it_relu_bwd": [vp, vp, C.c_long, vp],
            "it_loss_grad": [vp, vp, vp, vp, vp, i, C.c_float, vp],
            "it_adam": [vp, vp, vp, vp, i, C.c_float, C.c_float, C.c_float, C.c_float, i, C.c_float, vp],
        }.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        self.L = L

    def __getattr__(self, k):
        f = getattr(self.L, k)

        def call(*a):
            rc = f(*a)
            if rc != 0:
                raise RuntimeError(f"{k} failed ({rc})")
        return call


def train(feat, cell, cost, *, epochs=8, batch=512, lr=2e-3, cost_w=0.05, seed=0, params=None, log=None):
    """Trains the CNN with the HIP FP32 kernels on the current device; returns float32 parameters."""
    import torch

    L = _TrainLib()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    P = torch.tensor(init_params(seed) if params is None else params, device=dev)
    Gr = torch.zeros_like(P)
    M = torch.zeros_like(P)
    V = torch.zeros_like(P)
    sl = param_slices()
    X = torch.tensor(feat, device=dev)
    TC = torch.tensor(cell, device=dev, dtype=torch.int32)
    TK = torch.tensor(cost, device=dev)
    n = X.shape[0]
    acts = [torch.empty((batch, 81, cout), device=dev) for _, cout in LAYERS]
    grads = [torch.empty((batch, 81, cout), device=dev) for _, cout in LAYERS]
    dx = torch.empty((batch, 81, 32), device=dev)
    loss = torch.zeros(1, device=dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    step = 0
    ptr = lambda t: C.c_void_p(t.data_ptr())
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(dev)
        tot, nb = 0.0, 0
        for s in range(0, n - batch + 1, batch):
            idx = perm[s:s + batch]
            xb, tc, tk = X[idx].contiguous(), TC[idx].contiguous(), TK[idx].contiguous()
            inp = xb
            for li, ((w, b), (cin, cout)) in enumerate(zip(sl, LAYERS)):
                L.it_conv_fwd(ptr(inp), ptr(P[w[0]:]), ptr(P[b[0]:]), ptr(acts[li]), batch, cin, cout, int(li < 3), stream)
                inp = acts[li]
            grads[3].zero_()
            loss.zero_()
            L.it_loss_grad(ptr(acts[3]), ptr(tc), ptr(tk), ptr(grads[3]), ptr(loss), batch, cost_w, stream)
            Gr.zero_()
            for li in range(3, -1, -1):
                (w, b), (cin, cout) = sl[li], LAYERS[li]
                src = xb if li == 0 else acts[li - 1]
                L.it_conv_bwd_w(ptr(src), ptr(grads[li]), ptr(Gr[w[0]:]), ptr(Gr[b[0]:]), batch, cin, cout, stream)
                if li > 0:
                    L.it_conv_bwd_data(ptr(grads[li]), ptr(P[w[0]:]), ptr(grads[li - 1]), batch, cin, cout, stream)
                    L.it_relu_bwd(ptr(grads[li - 1]), ptr(acts[li - 1]), batch * 81 * cin, stream)
            step += 1
            L.it_adam(ptr(P), ptr(Gr), ptr(M), ptr(V), P.numel(), lr, 0.9, 0.999, 1e-8, step, 1e-5, stream)
            tot += float(loss.item())
            nb += 1
        if log:
            log(f"epoch {ep}: loss {tot / max(1, nb):.4f}")
    torch.cuda.synchronize()
    return P.cpu().numpy()


def main():
    """Builds the dataset from 300 generator clips (200 over the generator's range, 100 with noise
    sigma 3-5), trains, and writes ironav_amd/weights/proposal.npz (or the path in $IA_WEIGHTS_OUT)."""
    import time

    t0 = time.time()
    log = lambda m: print(f"[{time.time() - t0:7.1f}s] {m}", flush=True)
    # the generator's range of content, then as many again of heavy noise (sigma 3-5), where the cost
    # head would otherwise predict too low a SAD
    parts = [make_dataset(log=log), make_dataset(n_clips=100, seed=30000, noise=(3.0, 5.0), log=log)]
    feat, cell, cost = (np.concatenate(x) for x in zip(*parts))
    log(f"dataset: {len(feat)} samples, {np.mean(cell >= 0):.3f} of targets on the grid")
    params = train(feat, cell, cost, log=log)
    out = os.environ.get("IA_WEIGHTS_OUT", WEIGHTS)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.savez(out, params=params)
    log(f"wrote {out}")


if __name__ == "__main__":
    main()
