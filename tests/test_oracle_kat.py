"""Pins of the CPU oracle: known answers from the reference's own tests, closed forms and brute force.

The reference (nervosys/IronAV, Rust) cannot be built here, so each check below states what fixes the
expected value: a test of the reference restated with its inputs (file:line), a formula of the
standard, or an independent brute-force computation.
"""
import ctypes as C
import math

import numpy as np
import pytest

from ironav_amd import abi


def _lcg_luma(pw, ph, bd, seed=99):
    # crates/ironav-codec/src/h264/enc/subpel.rs:164-170
    out = np.empty(pw * ph, dtype=np.uint16)
    s = seed
    for i in range(pw * ph):
        s = (s * 1103515245 + 12345) & 0xFFFFFFFF
        out[i] = (s >> 16) % (1 << bd)
    return out.reshape(ph, pw)


@pytest.mark.parametrize("bd", [8, 10])
def test_halfpel_predict_equals_decoder_interpolation(oracle, bd):
    """enc/subpel.rs:160-193: HalfPel::predict equals the decoder's luma interpolation for a grid of
    vectors, including ones far outside the picture. The decoder side here is or_luma_mc_ref, a
    per-sample statement of H.264 8.4.2.2.1 with clamped fetches, independent of the planes."""
    pw, ph = 48, 32
    luma = _lcg_luma(pw, ph, bd)
    maxv = (1 << bd) - 1
    planes = oracle.halfpel_build(luma, bd)
    for x, y in [(0, 0), (16, 16), (32, 16)]:
        for my in range(-100, 100):
            for mx in [-97, -40, -3, -2, -1, 0, 1, 2, 3, 5, 6, 7, 30, 97]:
                a = oracle.halfpel_predict(planes, pw, ph, x, y, (mx, my), 16, 16)
                b = oracle.luma_mc_ref(luma, x, y, (mx, my), 16, 16, maxv)
                assert np.array_equal(a, b), (bd, x, y, mx, my)
    for x, y in [(0, 0), (40, 24), (32, 16)]:
        for my in [-2049, -1203, -300, -1, 2, 299, 1205, 2047]:
            for mx in [-8191, -2001, -401, 3, 398, 1999, 8190]:
                a = oracle.halfpel_predict(planes, pw, ph, x, y, (mx, my), 8, 8)
                b = oracle.luma_mc_ref(luma, x, y, (mx, my), 8, 8, maxv)
                assert np.array_equal(a, b), (bd, x, y, mx, my)


def test_halfpel_integer_plane_and_half_samples_closed_form(oracle):
    """G is the edge-extended picture; B and H are the 6-tap filter (1,-5,20,20,-5,1) rounded by 5 bits
    (H.264 8.4.2.2.1 eq. 8-243..8-246), checked with numpy on a random picture."""
    rng = np.random.default_rng(3)
    pw, ph = 32, 16
    luma = rng.integers(0, 256, size=(ph, pw)).astype(np.uint16)
    planes = oracle.halfpel_build(luma, 8).astype(np.int64)
    ext = np.pad(luma.astype(np.int64), 48, mode="edge")
    assert np.array_equal(planes[0], ext)
    taps = np.array([1, -5, 20, 20, -5, 1])
    sh, sw = ext.shape
    b = np.zeros_like(ext)
    for x in range(2, sw - 3):
        b[:, x] = np.clip((ext[:, x - 2:x + 4] @ taps + 16) >> 5, 0, 255)
    assert np.array_equal(planes[1], b)
    h = np.zeros_like(ext)
    for y in range(2, sh - 3):
        h[y] = np.clip((taps @ ext[y - 2:y + 4] + 16) >> 5, 0, 255)
    assert np.array_equal(planes[2], h)


def test_transform_round_trip_4x4(oracle):
    """enc/frame.rs:2443-2465: QP 0 quantisation then reconstruction returns the residual within 1."""
    L = oracle.lib()
    src = np.array([((i * 37) % 23) * 9 for i in range(16)], dtype=np.int32)
    d = (src - 100).astype(np.int32)
    c = np.zeros(16, dtype=np.int32)
    L.or_fdct4(d.ctypes.data, c.ctypes.data)
    mf0, v0 = [13107, 5243, 8066], [10, 16, 13]  # H.264 quantisation/level scale at QP%6 == 0

    def cls(p):
        a, b = p & 1, (p >> 2) & 1
        return 0 if (a, b) == (0, 0) else 1 if (a, b) == (1, 1) else 2

    co = np.array([L.or_quant(int(c[p]), mf0[cls(p)], 1 << 14, 15) * v0[cls(p)] for p in range(16)], dtype=np.int32)
    out = np.full(16, 100, dtype=np.uint16)
    L.or_idct4_add(co.ctypes.data, out.ctypes.data, 4, 255)
    assert np.all(np.abs(out.astype(int) - src) <= 1)


@pytest.mark.parametrize("qp", [0, 1, 5, 6])
def test_transform_round_trip_8x8(oracle, qp):
    """enc/frame.rs:2468-2479: 8x8 quantisation at low QP round-trips within 1 + qp/3."""
    L = oracle.lib()
    src = np.array([((i * 37) % 23) * 9 + (i // 8) * 5 for i in range(64)], dtype=np.int32)
    pred = np.full(64, 100, dtype=np.int32)
    lv, co = np.zeros(64, dtype=np.int32), np.zeros(64, dtype=np.int32)
    L.or_quant8(src.ctypes.data, pred.ctypes.data, qp, 1, lv.ctypes.data, co.ctypes.data)
    out = np.full(64, 100, dtype=np.uint16)
    L.or_idct8_add(co.ctypes.data, out.ctypes.data, 8, 255)
    assert np.all(np.abs(out.astype(int) - src) <= 1 + qp // 3)


def test_fdct4_is_the_h264_core_matrix(oracle):
    """The 4x4 forward core transform is Cf . X . Cf^T (H.264's integer DCT approximation)."""
    L = oracle.lib()
    cf = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]])
    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.integers(-255, 256, size=(4, 4)).astype(np.int32)
        out = np.zeros(16, dtype=np.int32)
        L.or_fdct4(np.ascontiguousarray(x).ctypes.data, out.ctypes.data)
        assert np.array_equal(out.reshape(4, 4), cf @ x @ cf.T)


def test_hadamard_and_satd_closed_forms(oracle):
    """hadamard4 is H.X.H with the (unnormalised) 4-point Hadamard in the order of h264/idct.rs:76-95;
    satd4 is half the sum of |H.D.H| (enc/frame.rs:144-161), order-independent."""
    L = oracle.lib()
    h = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]])
    rng = np.random.default_rng(6)
    for _ in range(50):
        x = rng.integers(-2000, 2000, size=(4, 4)).astype(np.int32)
        out = np.zeros(16, dtype=np.int32)
        L.or_hadamard4(np.ascontiguousarray(x).ctypes.data, out.ctypes.data)
        assert np.array_equal(out.reshape(4, 4), h @ x @ h.T)
        d = rng.integers(-255, 256, size=16).astype(np.int32)
        assert L.or_satd4(d.ctypes.data) == int(np.abs(h @ d.reshape(4, 4) @ h.T).sum()) >> 1


def test_decimation_known_answers(oracle):
    """enc/frame.rs:2482-2492."""
    L = oracle.lib()
    lv = np.zeros(16, dtype=np.int32)
    assert L.or_decimate_score(lv.ctypes.data) == 0
    lv[15] = 1
    assert L.or_decimate_score(lv.ctypes.data) == 0
    lv[15] = 0
    lv[0] = 1
    assert L.or_decimate_score(lv.ctypes.data) == 3
    lv[3] = -2
    assert L.or_decimate_score(lv.ctypes.data) == 9


def test_entropy_table_closed_form(oracle):
    """enc/cabac.rs:162-173: round(-log2 p * 256) with pLPS = 0.5 (0.01875/0.5)^(p/63)."""
    t = np.zeros(128, dtype=np.uint32)
    oracle.lib().or_entropy_table(t.ctypes.data)
    p = np.arange(64, dtype=np.float64)
    lps = 0.5 * (0.01875 / 0.5) ** (p / 63.0)
    mps = np.floor(-np.log2(1.0 - lps) * 256.0 + 0.5)
    lp = np.floor(-np.log2(lps) * 256.0 + 0.5)
    assert np.array_equal(t.reshape(64, 2)[:, 0], mps.astype(np.uint32))
    assert np.array_equal(t.reshape(64, 2)[:, 1], lp.astype(np.uint32))
    # the most probable state costs almost nothing, the least probable LPS about 5.7 bits
    assert t[0] == 256 and t[1] == 256 and t[126] < 8 and 1400 < t[127] < 1500


def test_exp_golomb_lengths(oracle):
    """enc/bitw.rs:81-89 against the length of the written codes: ue(v) is the binary of v+1 with
    len-1 leading zeros; se maps v>0 to 2v-1 and v<=0 to -2v."""
    L = oracle.lib()
    for v in range(0, 5000):
        assert L.or_ue_bits(v) == 2 * len(bin(v + 1)[2:]) - 1
    for v in range(-3000, 3000):
        k = 2 * v - 1 if v > 0 else -2 * v
        assert L.or_se_bits(v) == 2 * len(bin(k + 1)[2:]) - 1


def test_lambda_closed_form(oracle):
    """enc/frame.rs:164-166."""
    for qp in range(0, 52):
        assert oracle.lib().or_lambda(qp) == max(1, int(math.floor(2 ** ((qp - 12) / 6) + 0.5)))


def _np_fullsearch(src, ref, mb_w, mb_h, rng):
    pw, ph = mb_w * 16, mb_h * 16
    g = np.pad(ref.astype(np.int32), 48, mode="edge")
    out = np.zeros((mb_w * mb_h, abi.NPART, 3), dtype=np.int64)
    for mby in range(mb_h):
        for mbx in range(mb_w):
            x0, y0 = mbx * 16, mby * 16
            dxs = range(max(-rng, -(x0 + 24)), min(rng, pw - x0 + 8) + 1)
            dys = range(max(-rng, -(y0 + 24)), min(rng, ph - y0 + 8) + 1)
            s = src[y0:y0 + 16, x0:x0 + 16].astype(np.int32)
            best = [None] * abi.NPART
            for dy in dys:
                for dx in dxs:
                    r = g[y0 + dy + 48:y0 + dy + 64, x0 + dx + 48:x0 + dx + 64]
                    a = np.abs(s - r)
                    for p, (x, y, w, h) in enumerate(abi.PARTITIONS):
                        v = int(a[y * 4:(y + h) * 4, x * 4:(x + w) * 4].sum())
                        if best[p] is None or v < best[p][0]:
                            best[p] = (v, dx, dy)
            for p in range(abi.NPART):
                out[mby * mb_w + mbx, p] = best[p]
    return out


def test_fullsearch_brute_force(oracle):
    """The oracle's exhaustive search against an independent numpy brute force (window, partition
    numbering, first-minimum tie rule)."""
    rng = np.random.default_rng(11)
    mb_w, mb_h = 3, 2
    ref = rng.integers(0, 256, size=(32, 48)).astype(np.uint8)
    src = np.roll(ref, (3, -5), axis=(0, 1)).copy()
    src[5:9, 7:30] = rng.integers(0, 256, size=(4, 23))
    # flat areas produce ties that the first-minimum rule decides
    src[20:32, 0:20] = 77
    ref[0:20, 0:30] = 77
    fs = oracle.fullsearch(src, [ref], mb_w, mb_h, 6)
    want = _np_fullsearch(src, ref, mb_w, mb_h, 6)
    assert np.array_equal(fs[0]["sad"], want[:, :, 0])
    assert np.array_equal(fs[0]["mv"][:, :, 0], want[:, :, 1] * 4)
    assert np.array_equal(fs[0]["mv"][:, :, 1], want[:, :, 2] * 4)


def test_intra_16x16_and_chroma_closed_forms(oracle):
    """Vertical, horizontal and DC predictions are their definitions; plane prediction is
    Clip((a + b(x-7) + c(y-7) + 16) >> 5) with H.264 8.3.3.4's a, b, c computed here in numpy."""
    L = oracle.lib()
    rng = np.random.default_rng(8)
    for _ in range(20):
        top = rng.integers(0, 256, size=16).astype(np.int32)
        left = rng.integers(0, 256, size=16).astype(np.int32)
        tl = int(rng.integers(0, 256))
        out = np.zeros(256, dtype=np.uint16)
        for mode in range(4):
            L.or_predict_16x16(mode, top.ctypes.data, left.ctypes.data, tl, 1, 1, out.ctypes.data, 8)
            p = out.reshape(16, 16).astype(np.int64)
            if mode == 0:
                assert np.array_equal(p, np.tile(top, (16, 1)))
            elif mode == 1:
                assert np.array_equal(p, np.tile(left[:, None], (1, 16)))
            elif mode == 2:
                assert np.all(p == (top.sum() + left.sum() + 16) >> 5)
            else:
                tt = np.concatenate([[tl], top])  # tt[i + 1] = p[i, -1]
                ll = np.concatenate([[tl], left])
                H = sum((x + 1) * (tt[8 + x + 1] - tt[6 - x + 1]) for x in range(8))
                V = sum((y + 1) * (ll[8 + y + 1] - ll[6 - y + 1]) for y in range(8))
                a = 16 * (left[15] + top[15])
                b = (5 * H + 32) >> 6
                c = (5 * V + 32) >> 6
                yy, xx = np.mgrid[0:16, 0:16]
                assert np.array_equal(p, np.clip((a + b * (xx - 7) + c * (yy - 7) + 16) >> 5, 0, 255))
        # chroma DC: per-4x4 quadrant rules of 8.3.4.1-8.3.4.3
        ct = top[:8].copy()
        cl = left.copy()
        L.or_predict_chroma(0, ct.ctypes.data, cl.ctypes.data, tl, 1, 1, out.ctypes.data, 8, 8)
        p = out[:64].reshape(8, 8)
        assert p[0, 0] == (ct[:4].sum() + cl[:4].sum() + 4) >> 3
        assert p[0, 4] == (ct[4:8].sum() + 2) >> 2
        assert p[4, 0] == (cl[4:8].sum() + 2) >> 2
        assert p[4, 4] == (ct[4:8].sum() + cl[4:8].sum() + 4) >> 3


def test_intra_nxn_directional_symmetries(oracle):
    """Structural pins of the nine 4x4/8x8 modes: vertical/horizontal/DC are their definitions, every
    mode maps constant edges to the constant, diagonal-down-right is symmetric under transposition
    (swap top and left), and vertical-right is the transpose of horizontal-down with the edges swapped."""
    L = oracle.lib()
    rng = np.random.default_rng(9)
    for n in (4, 8):
        for _ in range(30):
            top = rng.integers(0, 256, size=16).astype(np.int32)
            left = np.zeros(8, dtype=np.int32)
            left[:n] = rng.integers(0, 256, size=n)
            tl = int(rng.integers(0, 256))

            def pred(mode, t, lft):
                out = np.zeros(n * n, dtype=np.uint16)
                t16 = np.zeros(16, dtype=np.int32)
                t16[:len(t)] = t
                l8 = np.zeros(8, dtype=np.int32)
                l8[:n] = lft[:n]
                L.or_predict_nxn(n, mode, t16.ctypes.data, l8.ctypes.data, tl, 1, 1, 1, out.ctypes.data, 8)
                return out.reshape(n, n).astype(np.int64)

            assert np.array_equal(pred(0, top, left), np.tile(top[:n], (n, 1)))
            assert np.array_equal(pred(1, top, left), np.tile(left[:n, None], (1, n)))
            assert np.all(pred(2, top, left) == (top[:n].sum() + left[:n].sum() + n) >> (2 if n == 4 else 3) + 1)
            assert np.array_equal(pred(4, top, left), pred(4, left, top[:n]).T)
            assert np.array_equal(pred(5, top, left), pred(6, left, top[:n]).T)
            const = np.full(16, 90, dtype=np.int32)
            for mode in range(9):
                out = np.zeros(n * n, dtype=np.uint16)
                L.or_predict_nxn(n, mode, const.ctypes.data, const[:8].ctypes.data, 90, 1, 1, 1, out.ctypes.data, 8)
                assert np.all(out == 90), (n, mode)


def test_abi_struct_layouts(oracle):
    """The ctypes mirror of include/ironav_h264a.h lays the structs out as the C compiler does."""
    sizes = np.zeros(6, dtype=np.int32)
    oracle.lib().or_abi_sizes(sizes.ctypes.data)
    assert list(sizes) == [C.sizeof(abi.Config), C.sizeof(abi.MbInfo), C.sizeof(abi.MbCode), C.sizeof(abi.MbStats),
                           C.sizeof(abi.FsResult), C.sizeof(abi.Frame)]
    assert abi.MBCODE_DTYPE.itemsize == C.sizeof(abi.MbCode)


def test_threaded_slices_equal_one_thread(oracle):
    """The picture driver with slices on host threads (the CPU baseline's form, one slice per thread as
    enc/mod.rs:623-627 runs them) codes every slice exactly as the single-threaded driver."""
    from ironav_amd import clipgen
    from ironav_amd.h264a import Picture

    w, h = 96, 64
    pics = [Picture(*clipgen.pad_planes(*f)) for f in clipgen.make_clip(clipgen.ClipSpec(w, h, 3, 41, max_motion=6))]
    cfg = abi.Config(w, h, 8, 1, 1, 0, 0, 2, 4, abi.SEARCH_DIAMOND, 16, 1.1)
    digests = []
    for threads in (1, 3):
        refs, out = [], []
        for t, pic in enumerate(pics):
            kind = abi.PIC_IDR if t == 0 else abi.PIC_P
            codes, stats, rec, mbs = oracle.analyse_picture(cfg, pic, 25 if t == 0 else 28, kind, refs[:2],
                                                            threads=threads)
            out.append((codes.tobytes(), stats.tobytes(), b"".join(r.tobytes() for r in rec), mbs.tobytes()))
            refs.insert(0, (Picture(*rec), mbs))
        digests.append(out)
    assert digests[0] == digests[1]
