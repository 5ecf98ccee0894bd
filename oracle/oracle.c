/* oracle.c — CPU restatement of nervosys/IronAV's H.264 per-macroblock analysis (P and I slices).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the GPU path. See oracle.h for what pins each part.
 * Build: make -C oracle (gcc, -ffp-contract=off so the f64 decision arithmetic is not fused).
 *
 * Paths below are relative to the IronAV checkout: enc/ = crates/ironav-codec/src/h264/enc/,
 * h264/ = crates/ironav-codec/src/h264/.
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "h264_tables.h"

enum { K_NONE = 0, K_I4 = 1, K_I8 = 2, K_I16 = 3, K_PCM = 4, K_INTER = 5 };

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int mini(int a, int b) { return a < b ? a : b; }
static inline int maxi(int a, int b) { return a > b ? a : b; }
static inline int absi(int a) { return a < 0 ? -a : a; }
/* v << s on the two's-complement bits, as Rust's i32 `<<` shifts negative dequantised levels */
static inline int shl(int v, int s) { return (int)((uint32_t)v << s); }

/* ---------------------------------------------------------------------------------------------
 * Bit lengths and cost helpers (enc/bitw.rs:81-89, enc/frame.rs:164-171)
 * ------------------------------------------------------------------------------------------- */

uint32_t or_ue_bits(uint32_t v) {
  uint32_t x = v + 1, len = 0;
  while (x) { len++; x >>= 1; }
  return 2 * len - 1;
}

uint32_t or_se_bits(int32_t v) {
  return or_ue_bits(v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * (int64_t)v));
}

/* lambda (enc/frame.rs:164-166): max(1, round(2^((qp-12)/6))) in f64 */
int32_t or_lambda(int qp) {
  int l = (int)round(pow(2.0, ((double)qp - 12.0) / 6.0));
  return l < 1 ? 1 : l;
}

static int mvd_bits(const int16_t *mv, const int16_t *mvp) {
  return (int)(or_se_bits((int)mv[0] - mvp[0]) + or_se_bits((int)mv[1] - mvp[1]));
}

static int16_t median3(int16_t a, int16_t b, int16_t c) {
  int16_t mx = a > b ? a : b, mn = a < b ? a : b;
  int16_t t = mn > c ? mn : c;
  return mx < t ? mx : t;
}

/* ---------------------------------------------------------------------------------------------
 * Half-sample planes (enc/subpel.rs:21-151)
 * ------------------------------------------------------------------------------------------- */

static int tap6(int a, int b, int c, int d, int e, int f) { return a - 5 * b + 20 * c + 20 * d - 5 * e + f; }

void or_halfpel_build(const uint16_t *luma, int pw, int ph, int bd, uint16_t *planes) {
  const int maxv = (1 << bd) - 1, sw = pw + 2 * OR_PAD, sh = ph + 2 * OR_PAD;
  const size_t n = (size_t)sw * sh;
  uint16_t *g = planes, *b = planes + n, *h = planes + 2 * n, *j = planes + 3 * n;
  int32_t *b1 = calloc(n, sizeof *b1);
  memset(planes, 0, 4 * n * sizeof *planes);
  /* enc/subpel.rs:70-83: edge-extended integer rows, and the horizontal 6-tap sums for x in 2..sw-3 */
  for (int y = 0; y < sh; y++) {
    int sy = mini(y > OR_PAD ? y - OR_PAD : 0, ph - 1);
    uint16_t *gr = g + (size_t)y * sw;
    for (int x = 0; x < sw; x++) gr[x] = luma[(size_t)sy * pw + mini(x > OR_PAD ? x - OR_PAD : 0, pw - 1)];
    for (int x = 2; x < sw - 3; x++) {
      int v = tap6(gr[x - 2], gr[x - 1], gr[x], gr[x + 1], gr[x + 2], gr[x + 3]);
      b1[(size_t)y * sw + x] = v;
      b[(size_t)y * sw + x] = (uint16_t)clampi((v + 16) >> 5, 0, maxv);
    }
  }
  /* enc/subpel.rs:91-107: vertical half samples and the diagonal ones from the unrounded sums */
  for (int y = 2; y < sh - 3; y++) {
    for (int x = 0; x < sw; x++) {
#define G_(yy) ((int)g[(size_t)(yy) * sw + x])
#define B1_(yy) (b1[(size_t)(yy) * sw + x])
      int v = tap6(G_(y - 2), G_(y - 1), G_(y), G_(y + 1), G_(y + 2), G_(y + 3));
      h[(size_t)y * sw + x] = (uint16_t)clampi((v + 16) >> 5, 0, maxv);
      v = tap6(B1_(y - 2), B1_(y - 1), B1_(y), B1_(y + 1), B1_(y + 2), B1_(y + 3));
      j[(size_t)y * sw + x] = (uint16_t)clampi((v + 512) >> 10, 0, maxv);
#undef G_
#undef B1_
    }
  }
  free(b1);
}

/* The two planes (and offsets) a quarter-sample fraction averages, indexed [fy][fx]:
 * (plane a, dx, dy, plane b, dx, dy) with planes 0 G, 1 B, 2 H, 3 J — H.264 8.4.2.2.1's
 * quarter samples as means of the nearest integer/half samples (enc/subpel.rs:26-38). */
static const uint8_t QMIX[4][4][6] = {
    {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 0}, {1, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 0, 0}},
    {{0, 0, 0, 2, 0, 0}, {1, 0, 0, 2, 0, 0}, {1, 0, 0, 3, 0, 0}, {1, 0, 0, 2, 1, 0}},
    {{2, 0, 0, 2, 0, 0}, {2, 0, 0, 3, 0, 0}, {3, 0, 0, 3, 0, 0}, {3, 0, 0, 2, 1, 0}},
    {{0, 0, 1, 2, 0, 0}, {2, 0, 0, 1, 0, 1}, {3, 0, 0, 1, 0, 1}, {2, 1, 0, 1, 0, 1}},
};

/* enc/subpel.rs:122-145 */
void or_halfpel_predict(const uint16_t *planes, int pw, int ph, int x, int y, int mvx, int mvy, int w, int h,
                        uint16_t *out) {
  const int sw = pw + 2 * OR_PAD, sh = ph + 2 * OR_PAD;
  const size_t n = (size_t)sw * sh;
  int fx = mvx & 3, fy = mvy & 3;
  int ix = clampi(x + (mvx >> 2), -w - 3, pw + 1);
  int iy = clampi(y + (mvy >> 2), -h - 3, ph + 1);
  const uint8_t *m = QMIX[fy][fx];
  const uint16_t *a = planes + m[0] * n + (size_t)(iy + m[2] + OR_PAD) * sw + (ix + m[1] + OR_PAD);
  if (m[0] == m[3] && m[1] == m[4] && m[2] == m[5]) {
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) out[r * w + c] = a[(size_t)r * sw + c];
  } else {
    const uint16_t *b = planes + m[3] * n + (size_t)(iy + m[5] + OR_PAD) * sw + (ix + m[4] + OR_PAD);
    for (int r = 0; r < h; r++)
      for (int c = 0; c < w; c++) out[r * w + c] = (uint16_t)((a[(size_t)r * sw + c] + b[(size_t)r * sw + c] + 1) >> 1);
  }
}

/* The decoder's interpolation, sample by sample from H.264 8.4.2.2.1 with reference samples fetched
 * at clamped coordinates (the decoder's edge rule, h264/inter.rs:466-547). Independent of the planes. */
void or_luma_mc_ref(const uint16_t *refp, int stride, int pw, int ph, int x, int y, int mx, int my, int w, int h,
                    uint16_t *out, int maxv) {
  int fx = mx & 3, fy = my & 3, x0 = x + (mx >> 2), y0 = y + (my >> 2);
#define G(xx, yy) ((int)refp[(size_t)clampi((yy), 0, ph - 1) * stride + clampi((xx), 0, pw - 1)])
#define B1(xx, yy) tap6(G((xx) - 2, yy), G((xx) - 1, yy), G(xx, yy), G((xx) + 1, yy), G((xx) + 2, yy), G((xx) + 3, yy))
#define BH(xx, yy) clampi((B1(xx, yy) + 16) >> 5, 0, maxv)
#define HV(xx, yy) clampi((tap6(G(xx, (yy) - 2), G(xx, (yy) - 1), G(xx, yy), G(xx, (yy) + 1), G(xx, (yy) + 2), G(xx, (yy) + 3)) + 16) >> 5, 0, maxv)
#define JJ(xx, yy) clampi((tap6(B1(xx, (yy) - 2), B1(xx, (yy) - 1), B1(xx, yy), B1(xx, (yy) + 1), B1(xx, (yy) + 2), B1(xx, (yy) + 3)) + 512) >> 10, 0, maxv)
  for (int r = 0; r < h; r++) {
    for (int c = 0; c < w; c++) {
      int X = x0 + c, Y = y0 + r, v;
      switch (fy * 4 + fx) {
        case 0: v = G(X, Y); break;
        case 1: v = (G(X, Y) + BH(X, Y) + 1) >> 1; break;
        case 2: v = BH(X, Y); break;
        case 3: v = (G(X + 1, Y) + BH(X, Y) + 1) >> 1; break;
        case 4: v = (G(X, Y) + HV(X, Y) + 1) >> 1; break;
        case 5: v = (BH(X, Y) + HV(X, Y) + 1) >> 1; break;
        case 6: v = (BH(X, Y) + JJ(X, Y) + 1) >> 1; break;
        case 7: v = (BH(X, Y) + HV(X + 1, Y) + 1) >> 1; break;
        case 8: v = HV(X, Y); break;
        case 9: v = (HV(X, Y) + JJ(X, Y) + 1) >> 1; break;
        case 10: v = JJ(X, Y); break;
        case 11: v = (JJ(X, Y) + HV(X + 1, Y) + 1) >> 1; break;
        case 12: v = (G(X, Y + 1) + HV(X, Y) + 1) >> 1; break;
        case 13: v = (HV(X, Y) + BH(X, Y + 1) + 1) >> 1; break;
        case 14: v = (JJ(X, Y) + BH(X, Y + 1) + 1) >> 1; break;
        default: v = (HV(X + 1, Y) + BH(X, Y + 1) + 1) >> 1; break;
      }
      out[r * w + c] = (uint16_t)v;
    }
  }
#undef G
#undef B1
#undef BH
#undef HV
#undef JJ
}

/* h264/inter.rs:377-386, 404-422: eighth-sample bilinear from a window clamped to the plane */
void or_chroma_mc(const uint16_t *refp, int pw, int ph, int x, int y, int mx, int my, int w, int h, uint16_t *out) {
  int fx = mx & 7, fy = my & 7, x0 = x + (mx >> 3), y0 = y + (my >> 3);
  int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
#define S(xx, yy) ((int)refp[(size_t)clampi((yy), 0, ph - 1) * pw + clampi((xx), 0, pw - 1)])
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      int X = x0 + c, Y = y0 + r;
      out[r * w + c] = (uint16_t)((wa * S(X, Y) + wb * S(X + 1, Y) + wc * S(X, Y + 1) + wd * S(X + 1, Y + 1) + 32) >> 6);
    }
#undef S
}

/* ---------------------------------------------------------------------------------------------
 * Transforms and quantisation (enc/frame.rs:20-161, h264/idct.rs)
 * ------------------------------------------------------------------------------------------- */

static int class4(int i) {
  int a = i & 1, b = (i >> 2) & 1;
  if (a == 0 && b == 0) return 0;
  if (a == 1 && b == 1) return 1;
  return 2;
}

static int class8(int p) {
  int i = p / 8, j = p % 8;
  if (i % 4 == 0 && j % 4 == 0) return 0;
  if (i % 2 == 1 && j % 2 == 1) return 1;
  if (i % 4 == 2 && j % 4 == 2) return 2;
  if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) return 3;
  if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) return 4;
  return 5;
}

static void fdct8_1d(const int32_t *d, int32_t *o) {
  int s07 = d[0] + d[7], s16 = d[1] + d[6], s25 = d[2] + d[5], s34 = d[3] + d[4];
  int a0 = s07 + s34, a1 = s16 + s25, a2 = s07 - s34, a3 = s16 - s25;
  int d07 = d[0] - d[7], d16 = d[1] - d[6], d25 = d[2] - d[5], d34 = d[3] - d[4];
  int a4 = d16 + d25 + (d07 + (d07 >> 1));
  int a5 = d07 - d34 - (d25 + (d25 >> 1));
  int a6 = d07 + d34 - (d16 + (d16 >> 1));
  int a7 = d16 - d25 + (d34 + (d34 >> 1));
  o[0] = a0 + a1;
  o[1] = a4 + (a7 >> 2);
  o[2] = a2 + (a3 >> 1);
  o[3] = a5 + (a6 >> 2);
  o[4] = a0 - a1;
  o[5] = a6 - (a5 >> 2);
  o[6] = (a2 >> 1) - a3;
  o[7] = (a4 >> 2) - a7;
}

void or_fdct8(const int32_t *d, int32_t *out) {
  int32_t t[64], col[8], v[8];
  for (int r = 0; r < 8; r++) fdct8_1d(d + r * 8, t + r * 8);
  for (int c = 0; c < 8; c++) {
    for (int r = 0; r < 8; r++) col[r] = t[r * 8 + c];
    fdct8_1d(col, v);
    for (int r = 0; r < 8; r++) out[r * 8 + c] = v[r];
  }
}

void or_fdct4(const int32_t *d, int32_t *o) {
  int32_t t[16];
  for (int r = 0; r < 4; r++) {
    const int32_t *a = d + r * 4;
    int s03 = a[0] + a[3], d03 = a[0] - a[3], s12 = a[1] + a[2], d12 = a[1] - a[2];
    t[r * 4] = s03 + s12;
    t[r * 4 + 1] = 2 * d03 + d12;
    t[r * 4 + 2] = s03 - s12;
    t[r * 4 + 3] = d03 - 2 * d12;
  }
  for (int c = 0; c < 4; c++) {
    int s03 = t[c] + t[12 + c], d03 = t[c] - t[12 + c], s12 = t[4 + c] + t[8 + c], d12 = t[4 + c] - t[8 + c];
    o[c] = s03 + s12;
    o[4 + c] = 2 * d03 + d12;
    o[8 + c] = s03 - s12;
    o[12 + c] = d03 - 2 * d12;
  }
}

int32_t or_quant(int32_t c, int32_t mf, int32_t f, int qbits) {
  int64_t a = ((int64_t)(c < 0 ? -(int64_t)c : c) * mf + f) >> qbits;
  return c < 0 ? -(int32_t)a : (int32_t)a;
}

/* enc/frame.rs:88-105 */
void or_quant8(const int32_t *src, const int32_t *pred, int qp, int intra, int32_t *lv, int32_t *co) {
  int32_t d[64], c[64];
  for (int k = 0; k < 64; k++) d[k] = src[k] - pred[k];
  or_fdct8(d, c);
  int m = qp % 6, qbits = 16 + qp / 6;
  int f = (1 << qbits) / (intra ? 3 : 6);
  memset(co, 0, 64 * sizeof *co);
  for (int k = 0; k < 64; k++) {
    int p = OR_ZIGZAG_DIRECT[k];
    int l = or_quant(c[p], OR_MF8[m][class8(p)], f, qbits);
    lv[k] = l;
    if (l != 0) {
      int v = l * 16 * OR_DEQUANT8_COEFF_INIT[m][class8(p)];
      co[p] = qp >= 36 ? shl(v, qp / 6 - 6) : (v + (1 << (5 - qp / 6))) >> (6 - qp / 6);
    }
  }
}

/* enc/frame.rs:144-161 */
int32_t or_satd4(const int32_t *d) {
  int32_t t[16];
  for (int r = 0; r < 4; r++) {
    int a = d[r * 4], b = d[r * 4 + 1], c = d[r * 4 + 2], e = d[r * 4 + 3];
    int s0 = a + b, s1 = c + e, d0 = a - b, d1 = c - e;
    t[r * 4] = s0 + s1;
    t[r * 4 + 1] = s0 - s1;
    t[r * 4 + 2] = d0 - d1;
    t[r * 4 + 3] = d0 + d1;
  }
  int sum = 0;
  for (int c = 0; c < 4; c++) {
    int a = t[c], b = t[4 + c], cc = t[8 + c], e = t[12 + c];
    int s0 = a + b, s1 = cc + e, d0 = a - b, d1 = cc - e;
    sum += absi(s0 + s1) + absi(s0 - s1) + absi(d0 - d1) + absi(d0 + d1);
  }
  return sum >> 1;
}

/* h264/idct.rs:8-32, with the wrapping i32 arithmetic of the reference */
#define WADD(a, b) ((int32_t)((uint32_t)(a) + (uint32_t)(b)))
#define WSUB(a, b) ((int32_t)((uint32_t)(a) - (uint32_t)(b)))
void or_idct4_add(const int32_t *c, uint16_t *dst, int stride, int maxv) {
  int32_t t[16];
  for (int i = 0; i < 4; i++) {
    const int32_t *d = c + i * 4;
    int32_t e0 = WADD(d[0], d[2]), e1 = WSUB(d[0], d[2]), e2 = WSUB(d[1] >> 1, d[3]), e3 = WADD(d[1], d[3] >> 1);
    t[i * 4] = WADD(e0, e3);
    t[i * 4 + 1] = WADD(e1, e2);
    t[i * 4 + 2] = WSUB(e1, e2);
    t[i * 4 + 3] = WSUB(e0, e3);
  }
  for (int j = 0; j < 4; j++) {
    int32_t g0 = WADD(t[j], t[8 + j]), g1 = WSUB(t[j], t[8 + j]), g2 = WSUB(t[4 + j] >> 1, t[12 + j]),
            g3 = WADD(t[4 + j], t[12 + j] >> 1);
    int32_t hh[4] = {WADD(g0, g3), WADD(g1, g2), WSUB(g1, g2), WSUB(g0, g3)};
    for (int i = 0; i < 4; i++) {
      uint16_t *p = dst + (size_t)i * stride + j;
      *p = (uint16_t)clampi(*p + (WADD(hh[i], 32) >> 6), 0, maxv);
    }
  }
}

static void row8(const int32_t *d, int32_t *o) {
  int32_t a0 = WADD(d[0], d[4]), a4 = WSUB(d[0], d[4]), a2 = WSUB(d[2] >> 1, d[6]), a6 = WADD(d[2], d[6] >> 1);
  int32_t b0 = WADD(a0, a6), b2 = WADD(a4, a2), b4 = WSUB(a4, a2), b6 = WSUB(a0, a6);
  int32_t a1 = WSUB(WSUB(WSUB(d[5], d[3]), d[7]), d[7] >> 1);
  int32_t a3 = WSUB(WSUB(WADD(d[1], d[7]), d[3]), d[3] >> 1);
  int32_t a5 = WADD(WADD(WSUB(d[7], d[1]), d[5]), d[5] >> 1);
  int32_t a7 = WADD(WADD(WADD(d[3], d[5]), d[1]), d[1] >> 1);
  int32_t b1 = WADD(a1, a7 >> 2), b7 = WSUB(a7, a1 >> 2), b3 = WADD(a3, a5 >> 2), b5 = WSUB(a3 >> 2, a5);
  o[0] = WADD(b0, b7);
  o[1] = WADD(b2, b5);
  o[2] = WADD(b4, b3);
  o[3] = WADD(b6, b1);
  o[4] = WSUB(b6, b1);
  o[5] = WSUB(b4, b3);
  o[6] = WSUB(b2, b5);
  o[7] = WSUB(b0, b7);
}

/* h264/idct.rs:55-74 */
void or_idct8_add(const int32_t *c, uint16_t *dst, int stride, int maxv) {
  int32_t t[64], d[8], hh[8];
  for (int i = 0; i < 8; i++) row8(c + i * 8, t + i * 8);
  for (int j = 0; j < 8; j++) {
    for (int i = 0; i < 8; i++) d[i] = t[i * 8 + j];
    row8(d, hh);
    for (int i = 0; i < 8; i++) {
      uint16_t *p = dst + (size_t)i * stride + j;
      *p = (uint16_t)clampi(*p + (WADD(hh[i], 32) >> 6), 0, maxv);
    }
  }
}

/* h264/idct.rs:76-95 */
void or_hadamard4(const int32_t *c, int32_t *o) {
  int32_t t[16];
  for (int i = 0; i < 4; i++) {
    const int32_t *d = c + i * 4;
    int32_t s0 = WADD(d[0], d[1]), s1 = WSUB(d[0], d[1]), s2 = WADD(d[2], d[3]), s3 = WSUB(d[2], d[3]);
    t[i * 4] = WADD(s0, s2);
    t[i * 4 + 1] = WSUB(s0, s2);
    t[i * 4 + 2] = WSUB(s1, s3);
    t[i * 4 + 3] = WADD(s1, s3);
  }
  for (int j = 0; j < 4; j++) {
    int32_t s0 = WADD(t[j], t[4 + j]), s1 = WSUB(t[j], t[4 + j]), s2 = WADD(t[8 + j], t[12 + j]), s3 = WSUB(t[8 + j], t[12 + j]);
    o[j] = WADD(s0, s2);
    o[4 + j] = WSUB(s0, s2);
    o[8 + j] = WSUB(s1, s3);
    o[12 + j] = WADD(s1, s3);
  }
}

/* enc/frame.rs:2369-2391 */
int32_t or_decimate_score(const int32_t *lv) {
  static const int RUN[16] = {3, 2, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int score = 0, i = 15;
  while (i >= 0 && lv[i] == 0) i--;
  while (i >= 0) {
    if (absi(lv[i]) > 1) return 9;
    i--;
    int run = 0;
    while (i >= 0 && lv[i] == 0) { i--; run++; }
    score += RUN[mini(run, 15)];
  }
  return score;
}

/* enc/frame.rs:2415-2436 */
int32_t or_decimate_score8(const int32_t *lv) {
  static const int RUN[25] = {3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0};
  int score = 0, i = 63;
  while (i >= 0 && lv[i] == 0) i--;
  while (i >= 0) {
    if (absi(lv[i]) > 1) return 9;
    i--;
    int run = 0;
    while (i >= 0 && lv[i] == 0) { i--; run++; }
    score += RUN[mini(run, 24)];
  }
  return score;
}

/* ---------------------------------------------------------------------------------------------
 * Intra prediction (h264/intra.rs)
 * ------------------------------------------------------------------------------------------- */

typedef struct {
  int32_t top[16], left[8], topleft;
  int have_top, have_left, have_topleft;
} edges_t;

static void predict_nxn(int n, int mode, const edges_t *e, uint16_t *out, int bd) {
#define T(i) ((i) < 0 ? e->topleft : e->top[(i)])
#define L(j) ((j) < 0 ? e->topleft : e->left[(j)])
  int shift = n == 4 ? 2 : 3;
  for (int y = 0; y < n; y++) {
    for (int x = 0; x < n; x++) {
      int v;
      switch (mode) {
        case 0: v = T(x); break;
        case 1: v = L(y); break;
        case 2: {
          int st = 0, sl = 0;
          for (int i = 0; i < n; i++) { st += T(i); sl += L(i); }
          if (e->have_top && e->have_left) v = (st + sl + n) >> (shift + 1);
          else if (e->have_left) v = (sl + n / 2) >> shift;
          else if (e->have_top) v = (st + n / 2) >> shift;
          else v = 1 << (bd - 1);
          break;
        }
        case 3:
          if (x == n - 1 && y == n - 1) v = (T(2 * n - 2) + 3 * T(2 * n - 1) + 2) >> 2;
          else v = (T(x + y) + 2 * T(x + y + 1) + T(x + y + 2) + 2) >> 2;
          break;
        case 4:
          if (x > y) v = (T(x - y - 2) + 2 * T(x - y - 1) + T(x - y) + 2) >> 2;
          else if (x < y) v = (L(y - x - 2) + 2 * L(y - x - 1) + L(y - x) + 2) >> 2;
          else v = (T(0) + 2 * e->topleft + L(0) + 2) >> 2;
          break;
        case 5: {
          int z = 2 * x - y;
          if (z >= 0 && z % 2 == 0) v = (T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 1) >> 1;
          else if (z >= 0) v = (T(x - (y >> 1) - 2) + 2 * T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 2) >> 2;
          else if (z == -1) v = (L(0) + 2 * e->topleft + T(0) + 2) >> 2;
          else v = (L(y - 2 * x - 1) + 2 * L(y - 2 * x - 2) + L(y - 2 * x - 3) + 2) >> 2;
          break;
        }
        case 6: {
          int z = 2 * y - x;
          if (z >= 0 && z % 2 == 0) v = (L(y - (x >> 1) - 1) + L(y - (x >> 1)) + 1) >> 1;
          else if (z >= 0) v = (L(y - (x >> 1) - 2) + 2 * L(y - (x >> 1) - 1) + L(y - (x >> 1)) + 2) >> 2;
          else if (z == -1) v = (L(0) + 2 * e->topleft + T(0) + 2) >> 2;
          else v = (T(x - 2 * y - 1) + 2 * T(x - 2 * y - 2) + T(x - 2 * y - 3) + 2) >> 2;
          break;
        }
        case 7:
          if (y % 2 == 0) v = (T(x + (y >> 1)) + T(x + (y >> 1) + 1) + 1) >> 1;
          else v = (T(x + (y >> 1)) + 2 * T(x + (y >> 1) + 1) + T(x + (y >> 1) + 2) + 2) >> 2;
          break;
        default: {
          int z = x + 2 * y, edge = 2 * n - 3;
          if (z > edge) v = L(n - 1);
          else if (z == edge) v = (L(n - 2) + 3 * L(n - 1) + 2) >> 2;
          else if (z % 2 == 0) v = (L(y + (x >> 1)) + L(y + (x >> 1) + 1) + 1) >> 1;
          else v = (L(y + (x >> 1)) + 2 * L(y + (x >> 1) + 1) + L(y + (x >> 1) + 2) + 2) >> 2;
          break;
        }
      }
      out[y * n + x] = (uint16_t)v;
    }
  }
#undef T
#undef L
}

void or_predict_nxn(int n, int mode, const int32_t *top16, const int32_t *left8, int32_t topleft, int have_top,
                    int have_left, int have_topleft, uint16_t *out, int bd) {
  edges_t e;
  memcpy(e.top, top16, sizeof e.top);
  memcpy(e.left, left8, sizeof e.left);
  e.topleft = topleft;
  e.have_top = have_top;
  e.have_left = have_left;
  e.have_topleft = have_topleft;
  predict_nxn(n, mode, &e, out, bd);
}

/* h264/intra.rs:130-159 */
static void filter_8x8_edges(edges_t *e) {
  edges_t p = *e;
  if (p.have_top) {
    e->top[0] = p.have_topleft ? (p.topleft + 2 * p.top[0] + p.top[1] + 2) >> 2 : (3 * p.top[0] + p.top[1] + 2) >> 2;
    for (int x = 1; x < 15; x++) e->top[x] = (p.top[x - 1] + 2 * p.top[x] + p.top[x + 1] + 2) >> 2;
    e->top[15] = (p.top[14] + 3 * p.top[15] + 2) >> 2;
  }
  if (p.have_topleft) {
    if (p.have_top && p.have_left) e->topleft = (p.top[0] + 2 * p.topleft + p.left[0] + 2) >> 2;
    else if (p.have_top) e->topleft = (3 * p.topleft + p.top[0] + 2) >> 2;
    else if (p.have_left) e->topleft = (3 * p.topleft + p.left[0] + 2) >> 2;
  }
  if (p.have_left) {
    e->left[0] = p.have_topleft ? (p.topleft + 2 * p.left[0] + p.left[1] + 2) >> 2 : (3 * p.left[0] + p.left[1] + 2) >> 2;
    for (int y = 1; y < 7; y++) e->left[y] = (p.left[y - 1] + 2 * p.left[y] + p.left[y + 1] + 2) >> 2;
    e->left[7] = (p.left[6] + 3 * p.left[7] + 2) >> 2;
  }
}

void or_filter_8x8_edges(int32_t *top16, int32_t *left8, int32_t *topleft, int have_top, int have_left,
                         int have_topleft) {
  edges_t e;
  memcpy(e.top, top16, sizeof e.top);
  memcpy(e.left, left8, sizeof e.left);
  e.topleft = *topleft;
  e.have_top = have_top;
  e.have_left = have_left;
  e.have_topleft = have_topleft;
  filter_8x8_edges(&e);
  memcpy(top16, e.top, sizeof e.top);
  memcpy(left8, e.left, sizeof e.left);
  *topleft = e.topleft;
}

/* h264/intra.rs:161-210 */
void or_predict_16x16(int mode, const int32_t *top, const int32_t *left, int32_t topleft, int have_top,
                      int have_left, uint16_t *out, int bd) {
  if (mode == 0) {
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) out[y * 16 + x] = (uint16_t)top[x];
  } else if (mode == 1) {
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) out[y * 16 + x] = (uint16_t)left[y];
  } else if (mode == 2) {
    int st = 0, sl = 0, v;
    for (int i = 0; i < 16; i++) { st += top[i]; sl += left[i]; }
    if (have_top && have_left) v = (st + sl + 16) >> 5;
    else if (have_left) v = (sl + 8) >> 4;
    else if (have_top) v = (st + 8) >> 4;
    else v = 1 << (bd - 1);
    for (int i = 0; i < 256; i++) out[i] = (uint16_t)v;
  } else {
#define TT(i) ((i) < 0 ? topleft : top[(i)])
#define LL(j) ((j) < 0 ? topleft : left[(j)])
    int h = 0, v = 0;
    for (int k = 0; k < 8; k++) {
      h += (k + 1) * (TT(8 + k) - TT(6 - k));
      v += (k + 1) * (LL(8 + k) - LL(6 - k));
    }
    int a = 16 * (left[15] + top[15]), b = (5 * h + 32) >> 6, c = (5 * v + 32) >> 6;
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) out[y * 16 + x] = (uint16_t)clampi((a + b * (x - 7) + c * (y - 7) + 16) >> 5, 0, (1 << bd) - 1);
#undef TT
#undef LL
  }
}

/* h264/intra.rs:212-287 */
void or_predict_chroma(int mode, const int32_t *top, const int32_t *left, int32_t topleft, int have_top,
                       int have_left, uint16_t *out, int bd, int h) {
  if (mode == 0) {
    for (int by = 0; by < h / 4; by++)
      for (int bx = 0; bx < 2; bx++) {
        int st = 0, sl = 0, v;
        for (int i = 0; i < 4; i++) { st += top[bx * 4 + i]; sl += left[by * 4 + i]; }
        if ((bx == 0 && by == 0) || (bx > 0 && by > 0)) {
          if (have_top && have_left) v = (st + sl + 4) >> 3;
          else if (have_left) v = (sl + 2) >> 2;
          else if (have_top) v = (st + 2) >> 2;
          else v = 1 << (bd - 1);
        } else if (bx > 0) {
          if (have_top) v = (st + 2) >> 2;
          else if (have_left) v = (sl + 2) >> 2;
          else v = 1 << (bd - 1);
        } else if (have_left) v = (sl + 2) >> 2;
        else if (have_top) v = (st + 2) >> 2;
        else v = 1 << (bd - 1);
        for (int y = 0; y < 4; y++)
          for (int x = 0; x < 4; x++) out[(by * 4 + y) * 8 + bx * 4 + x] = (uint16_t)v;
      }
  } else if (mode == 1) {
    for (int y = 0; y < h; y++)
      for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint16_t)left[y];
  } else if (mode == 2) {
    for (int y = 0; y < h; y++)
      for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint16_t)top[x];
  } else {
    int yc = (h - 8) / 2, hh = 0, vv = 0;
#define TT(i) ((i) < 0 ? topleft : top[(i)])
#define LL(j) ((j) < 0 ? topleft : left[(j)])
    for (int k = 0; k < 4; k++) hh += (k + 1) * (TT(4 + k) - TT(2 - k));
    for (int k = 0; k < 4 + yc; k++) vv += (k + 1) * (LL(4 + yc + k) - LL(2 + yc - k));
    int a = 16 * (left[h - 1] + top[7]), b = (34 * hh + 32) >> 6, c = ((h == 16 ? 5 : 34) * vv + 32) >> 6;
    for (int y = 0; y < h; y++)
      for (int x = 0; x < 8; x++) out[y * 8 + x] = (uint16_t)clampi((a + b * (x - 3) + c * (y - 3 - yc) + 16) >> 5, 0, (1 << bd) - 1);
#undef TT
#undef LL
  }
}

/* ---------------------------------------------------------------------------------------------
 * CABAC bit model (enc/cabac.rs:32-68, 150-246)
 * ------------------------------------------------------------------------------------------- */

static uint32_t ENTROPY[64][2];
static int entropy_ready;

/* enc/cabac.rs:162-173: -log2 of the MPS and LPS probabilities per pStateIdx, in 1/256 bits */
static void entropy_init(void) {
  if (entropy_ready) return;
  for (int p = 0; p < 64; p++) {
    double lps = 0.5 * pow(0.01875 / 0.5, (double)p / 63.0);
    ENTROPY[p][0] = (uint32_t)round(-log2(1.0 - lps) * 256.0);
    ENTROPY[p][1] = (uint32_t)round(-log2(lps) * 256.0);
  }
  entropy_ready = 1;
}

void or_entropy_table(uint32_t *out) {
  entropy_init();
  memcpy(out, ENTROPY, sizeof ENTROPY);
}

/* enc/cabac.rs:32-40 */
static uint8_t next_state(uint8_t s, int bin) {
  int p = s >> 1, mps = s & 1;
  if (bin != mps) {
    int nmps = p == 0 ? !mps : mps;
    return (uint8_t)((OR_TRANS_LPS[p] << 1) | nmps);
  }
  return (uint8_t)(((p < 62 ? p + 1 : p) << 1) | mps);
}

/* enc/cabac.rs:54-64: context states of a slice */
static void cabac_init(uint8_t *state, int intra, int slice_qp) {
  int qp = clampi(slice_qp, 0, 51);
  for (int i = 0; i < 1024; i++) {
    int m = intra ? OR_CABAC_INIT_I[i][0] : OR_CABAC_INIT_P[i][0];
    int n = intra ? OR_CABAC_INIT_I[i][1] : OR_CABAC_INIT_P[i][1];
    int pre = clampi(((m * qp) >> 4) + n, 1, 126);
    state[i] = pre <= 63 ? (uint8_t)((63 - pre) << 1) : (uint8_t)(((pre - 64) << 1) | 1);
  }
}

/* A bin sink over a set of context states: counts the bits `CabacCost` counts (enc/cabac.rs:176-191)
 * and adapts the states; used both on a copy (rate estimates) and on the live states (the writer's
 * context evolution, which depends on the bins alone). */
typedef struct {
  uint8_t *state;
  uint32_t bits;
} bins_t;

static void decision(bins_t *c, int ctx, int bin) {
  uint8_t s = c->state[ctx];
  c->bits += ENTROPY[s >> 1][bin != (s & 1)];
  c->state[ctx] = next_state(s, bin);
}
static void bypass(bins_t *c, int bin) {
  (void)bin;
  c->bits += 256;
}
static void terminate(bins_t *c, int bin) { c->bits += bin ? 7 * 256 : 1; }

/* enc/cabac.rs:17-29 */
static void exp_golomb(bins_t *c, uint32_t v, uint32_t k) {
  while (v >= (1u << k)) {
    bypass(c, 1);
    v -= 1u << k;
    k++;
  }
  bypass(c, 0);
  while (k > 0) {
    k--;
    bypass(c, (v >> k) & 1);
  }
}

/* enc/cabac.rs:194-246 */
static int write_block(bins_t *c, int cat, int cbf_ctx, const int32_t *levels, int max) {
  int total = 0, last = -1;
  for (int i = 0; i < max; i++)
    if (levels[i] != 0) { total++; last = i; }
  if (cbf_ctx >= 0) decision(c, cbf_ctx, total > 0);
  if (total == 0) return 0;
  for (int i = 0; i < max - 1; i++) {
    int sig, lst;
    if (cat == 3) { sig = mini(i, 2); lst = mini(i, 2); }
    else if (cat == 5) { sig = OR_SIG_OFF_8X8[0][i]; lst = OR_LAST_OFF_8X8[i]; }
    else { sig = i; lst = i; }
    int l = levels[i];
    decision(c, OR_SIG_OFF[0][cat] + sig, l != 0);
    if (l != 0) {
      decision(c, OR_LAST_OFF[0][cat] + lst, i == last);
      if (i == last) break;
    }
  }
  int base = OR_ABS_OFF[cat], eq1 = 0, gt1 = 0;
  for (int i = last; i >= 0; i--) {
    int l = levels[i];
    if (l == 0) continue;
    uint32_t a = (uint32_t)absi(l);
    int inc0 = gt1 != 0 ? 0 : mini(1 + eq1, 4);
    if (a == 1) {
      decision(c, base + inc0, 0);
      eq1++;
    } else {
      decision(c, base + inc0, 1);
      int incn = base + 5 + mini(gt1, 4 - (cat == 3));
      for (int k = 0; k < mini((int)a - 2, 13); k++) decision(c, incn, 1);
      if (a < 15) decision(c, incn, 0);
      else exp_golomb(c, a - 15, 0);
      gt1++;
    }
    bypass(c, l < 0);
  }
  return total;
}

/* enc/frame.rs:2393-2413 */
static void mvd_bins(bins_t *c, int comp, int d, uint32_t sum) {
  int base = comp == 0 ? 40 : 47;
  int inc = sum < 3 ? 0 : (sum > 32 ? 2 : 1);
  uint32_t a = (uint32_t)absi(d);
  decision(c, base + inc, a > 0);
  if (a > 0) {
    int ctx = base + 3;
    for (uint32_t k = 1; k < (a < 9 ? a : 9); k++) {
      decision(c, ctx, 1);
      if (ctx < base + 6) ctx++;
    }
    if (a < 9) decision(c, ctx, 0);
    else exp_golomb(c, a - 9, 3);
    bypass(c, d < 0);
  }
}

/* ---------------------------------------------------------------------------------------------
 * Slice coder (enc/frame.rs:336-2339, P and I slices)
 * ------------------------------------------------------------------------------------------- */

/* Partitions in 4x4 units (enc/frame.rs:183-184, 200-203) */
static const uint8_t PARTS[4][4][4] = {
    {{0, 0, 4, 4}}, {{0, 0, 4, 2}, {0, 2, 4, 2}}, {{0, 0, 2, 4}, {2, 0, 2, 4}},
    {{0, 0, 2, 2}, {2, 0, 2, 2}, {0, 2, 2, 2}, {2, 2, 2, 2}}};
static const int NPARTS[4] = {1, 2, 2, 4};
static const uint8_t SUBS[4][4][4] = {
    {{0, 0, 2, 2}}, {{0, 0, 2, 1}, {0, 1, 2, 1}}, {{0, 0, 1, 2}, {1, 0, 1, 2}},
    {{0, 0, 1, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}}};
static const int NSUBS[4] = {1, 2, 2, 4};

typedef struct { int x, y, w, h; } rect_t;

/* enc/frame.rs:207-212 */
static int rects(int part, const uint8_t *sub, rect_t *out) {
  int n = 0;
  if (part != 3) {
    for (int k = 0; k < NPARTS[part]; k++) out[n++] = (rect_t){PARTS[part][k][0], PARTS[part][k][1], PARTS[part][k][2], PARTS[part][k][3]};
    return n;
  }
  for (int k = 0; k < 4; k++)
    for (int s = 0; s < NSUBS[sub[k]]; s++) {
      const uint8_t *r = SUBS[sub[k]][s];
      out[n++] = (rect_t){(k & 1) * 2 + r[0], (k >> 1) * 2 + r[1], r[2], r[3]};
    }
  return n;
}

static int ref_of(int part, int n, int x, int y) { return part == 3 ? (y / 2) * 2 + x / 2 : n; }

typedef struct {
  uint8_t i4, inter, t8x8, skip, kind, cbp, chroma_mode, cbf_dc;
  uint8_t modes[16], nnz[16], cnnz[2][4];
  int16_t mvs[16][2], mvs1[16][2];   /* lists 0 and 1 */
  int8_t refs[4], refs1[4];
  uint8_t bdirect;                    /* B_Skip or B_Direct_16x16 */
  uint8_t mvd[16][2], mvd1[16][2];
} mbstate_t;

typedef struct {
  int kind, skip, i16_mode, chroma_mode, part, t8x8, cbp;
  uint8_t modes[16];
  int16_t mvs[16][2];
  int32_t mvds[16][2];
  uint8_t refs[4], sub[4];
  int32_t luma8[4][64];
  int32_t luma[16][16];
  int32_t dc[16];
  int32_t cdc[2][4];
  int32_t cac[2][4][16];
  uint8_t nnz[16], cnnz[2][4];
  /* B slices (enc/frame.rs:292-296): per partition the direction (0 direct, bit 0 list 0, bit 1
     list 1), per 8x8 block the lists used, per list the vectors per 4x4 block and the differences
     per partition */
  uint8_t bdir[4], buse[4][2];
  int16_t bmv[2][16][2];
  int32_t bmvd[2][4][2];
} mbcode_t;

static void mbcode_new(mbcode_t *c, int kind) {
  memset(c, 0, sizeof *c);
  c->kind = kind;
  memset(c->modes, 2, sizeof c->modes);
}

typedef struct {
  uint16_t *hp;          /* half-sample planes */
  const uint16_t *planes[3];
  const ia_mbinfo *mbs;  /* may be NULL: no inter macroblocks */
} refpic_t;

typedef struct {
  const uint16_t *src[3];
  int bd, mb_w, mb_h, qp, qpc, lambda;
  int row0, row1, slice_id;
  int p_slice, nref, sub8, use_cabac, transform8;
  /* a B slice: plist holds the list 0 and list 1 pictures, dist their picture-order distances
     (enc/frame.rs:350-354, with_b :442-447) */
  int b_slice, dist_tb, dist_td;
  const refpic_t *plist;
  uint16_t *pic[3];
  ia_mbinfo *mbs;
  mbstate_t *state;
  uint8_t cabac[1024];
  const ia_fsresult *fs;
  int search_mode;
  int fast;              /* realtime decisions, enc/frame.rs:348 */
  ia_mbstats *st;
} coder_t;

static int ls_(const coder_t *c) { return c->mb_w * 16; }
static int cs_(const coder_t *c) { return c->mb_w * 8; }
static int maxv_(const coder_t *c) { return (1 << c->bd) - 1; }
static int qpp_(const coder_t *c) { return c->qp + 6 * (c->bd - 8); }
static int up_(const coder_t *c, int mby) { return mby > c->row0; }

/* enc/frame.rs:509-515 */
static const mbstate_t *nb_state(const coder_t *c, int mbx, int mby) {
  if (mbx < 0 || mby < c->row0 || mbx >= c->mb_w || mby >= c->mb_h) return NULL;
  return &c->state[(mby - c->row0) * c->mb_w + mbx];
}

/* enc/frame.rs:521-539. Returns 1 when available, with *r and mv set. */
static int nb_motion(const coder_t *c, int l, int mbx, int mby, const int16_t (*cur)[2], const int8_t *cref, uint32_t done,
                     int x, int y, int *r, int16_t *mv) {
  int dx, dy, bx, by;
  if (x < 0 && y < 0) { dx = -1; dy = -1; bx = 3; by = 3; }
  else if (x < 0) { dx = -1; dy = 0; bx = 3; by = y; }
  else if (x >= 4 && y < 0) { dx = 1; dy = -1; bx = 0; by = 3; }
  else if (x >= 4) return 0;
  else if (y < 0) { dx = 0; dy = -1; bx = x; by = 3; }
  else {
    int b = y * 4 + x;
    if (!((done >> b) & 1)) return 0;
    if ((done >> (16 + b)) & 1) { *r = -1; mv[0] = mv[1] = 0; }
    else { *r = cref[b]; mv[0] = cur[b][0]; mv[1] = cur[b][1]; }
    return 1;
  }
  const mbstate_t *s = nb_state(c, mbx + dx, mby + dy);
  if (!s) return 0;
  int rr = (l ? s->refs1 : s->refs)[(by / 2) * 2 + bx / 2];
  const int16_t (*sm)[2] = l ? s->mvs1 : s->mvs;
  if (s->inter && rr >= 0) { *r = rr; mv[0] = sm[by * 4 + bx][0]; mv[1] = sm[by * 4 + bx][1]; }
  else { *r = -1; mv[0] = mv[1] = 0; }
  return 1;
}

/* enc/frame.rs:544-574: the vector predictor of the partition at (bx, by), w blocks wide (8.4.1.3) */
static void pred_mv(const coder_t *c, int l, int mbx, int mby, const int16_t (*cur)[2], const int8_t *cref, uint32_t done,
                    int bx, int by, int w, int dir, int rf, int16_t *out) {
  int ra = -1, rb = -1, rc = -1;
  int16_t ma[2] = {0, 0}, mb[2] = {0, 0}, mc[2] = {0, 0};
  int ha = nb_motion(c, l, mbx, mby, cur, cref, done, bx - 1, by, &ra, ma);
  int hb = nb_motion(c, l, mbx, mby, cur, cref, done, bx, by - 1, &rb, mb);
  int hc = nb_motion(c, l, mbx, mby, cur, cref, done, bx + w, by - 1, &rc, mc);
  if (!hc) hc = nb_motion(c, l, mbx, mby, cur, cref, done, bx - 1, by - 1, &rc, mc);
  if (!ha) { ra = -1; ma[0] = ma[1] = 0; }
  if (!hb) { rb = -1; mb[0] = mb[1] = 0; }
  if (!hc) { rc = -1; mc[0] = mc[1] = 0; }
#define RET(m) do { out[0] = (m)[0]; out[1] = (m)[1]; return; } while (0)
  if (dir == 1) {
    if (by == 0 && rb == rf) RET(mb);
    if (by != 0 && ra == rf) RET(ma);
  } else if (dir == 2) {
    if (bx == 0 && ra == rf) RET(ma);
    if (bx != 0 && rc == rf) RET(mc);
  }
  if (!hb && !hc && ha) RET(ma);
  int ea = ra == rf, eb = rb == rf, ec = rc == rf;
  if (ea + eb + ec == 1) {
    if (ea) RET(ma);
    if (eb) RET(mb);
    RET(mc);
  }
  out[0] = median3(ma[0], mb[0], mc[0]);
  out[1] = median3(ma[1], mb[1], mc[1]);
#undef RET
}

static const int16_t ZMV[16][2];
static const int8_t NOREF[16];

static void mv_pred(const coder_t *c, int l, int mbx, int mby, int rf, int16_t *out) {
  pred_mv(c, l, mbx, mby, ZMV, NOREF, 0, 0, 0, 4, 0, rf, out);
}

/* enc/frame.rs:582-595 */
static void part_mvps(const coder_t *c, int mbx, int mby, int part, const uint8_t *sub, const int16_t (*mvs)[2],
                      const uint8_t *refs, int16_t (*out)[2]) {
  int16_t cur[16][2] = {{0}};
  int8_t cref[16] = {0};
  uint32_t done = 0;
  rect_t rs[16];
  int n = rects(part, sub, rs);
  int dir = (int[]){0, 1, 2, 0}[part];
  for (int k = 0; k < n; k++) {
    int rf = refs[ref_of(part, k, rs[k].x, rs[k].y)];
    pred_mv(c, 0, mbx, mby, (const int16_t(*)[2])cur, cref, done, rs[k].x, rs[k].y, rs[k].w, dir, rf, out[k]);
    for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
      for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++) {
        cur[by * 4 + bx][0] = mvs[k][0];
        cur[by * 4 + bx][1] = mvs[k][1];
        cref[by * 4 + bx] = (int8_t)rf;
        done |= 1u << (by * 4 + bx);
      }
  }
}

/* enc/frame.rs:598-603 (8.4.1.1) */
static void skip_mv(const coder_t *c, int mbx, int mby, int16_t *out) {
  int ra, rb;
  int16_t ma[2], mb[2];
  int ha = nb_motion(c, 0, mbx, mby, ZMV, NOREF, 0, -1, 0, &ra, ma);
  int hb = nb_motion(c, 0, mbx, mby, ZMV, NOREF, 0, 0, -1, &rb, mb);
  if (!ha || !hb || (ra == 0 && ma[0] == 0 && ma[1] == 0) || (rb == 0 && mb[0] == 0 && mb[1] == 0)) {
    out[0] = out[1] = 0;
    return;
  }
  mv_pred(c, 0, mbx, mby, 0, out);
}

/* enc/frame.rs:619-627 */
static int pred_i4_mode(const coder_t *c, const uint8_t *cur, int mbx, int mby, int bx, int by) {
  int a = -1, b = -1;
  if (bx > 0) a = cur[by * 4 + bx - 1];
  else {
    const mbstate_t *s = nb_state(c, mbx - 1, mby);
    if (s) a = s->i4 ? s->modes[by * 4 + 3] : 2;
  }
  if (by > 0) b = cur[(by - 1) * 4 + bx];
  else {
    const mbstate_t *s = nb_state(c, mbx, mby - 1);
    if (s) b = s->i4 ? s->modes[12 + bx] : 2;
  }
  if (a >= 0 && b >= 0) return mini(a, b);
  return 2;
}

static void src_block(const coder_t *c, int plane, int x, int y, int n, int32_t *out) {
  int s = plane == 0 ? ls_(c) : cs_(c);
  for (int r = 0; r < n; r++)
    for (int k = 0; k < n; k++) out[r * n + k] = c->src[plane][(size_t)(y + r) * s + x + k];
}

/* enc/frame.rs:641-659 */
static int satd_block(const coder_t *c, int plane, int x, int y, int w, int h, const uint16_t *pred) {
  int s = plane == 0 ? ls_(c) : cs_(c), total = 0;
  const uint16_t *src = c->src[plane];
  for (int by = 0; by < h; by += 4)
    for (int bx = 0; bx < w; bx += 4) {
      int32_t d[16];
      for (int r = 0; r < 4; r++)
        for (int k = 0; k < 4; k++) d[r * 4 + k] = (int)src[(size_t)(y + by + r) * s + x + bx + k] - pred[(by + r) * w + bx + k];
      total += or_satd4(d);
    }
  return total;
}

/* The reference r of list l: a P slice's list 0 is plist, a B slice has one picture per list
 * (enc/frame.rs:677-702 ref_half / ref_pic). */
static const refpic_t *ref_pic(const coder_t *c, int l, int r) { return &c->plist[c->b_slice ? l : r]; }

/* enc/frame.rs:662-671 */
static int sad_int(const coder_t *c, int l, int r, int x, int y, int w, int h, const int16_t *mv) {
  const int pw = c->mb_w * 16, sw = pw + 2 * OR_PAD, ls = ls_(c);
  const uint16_t *g = ref_pic(c, l, r)->hp;
  int gx = x + (mv[0] >> 2), gy = y + (mv[1] >> 2), sad = 0;
  for (int rr = 0; rr < h; rr++)
    for (int k = 0; k < w; k++)
      sad += absi((int)c->src[0][(size_t)(y + rr) * ls + x + k] - g[(size_t)(gy + rr + OR_PAD) * sw + gx + k + OR_PAD]);
  return sad;
}

static void luma_pred_l(const coder_t *c, int l, int r, int x, int y, int w, int h, const int16_t *mv, uint16_t *out) {
  or_halfpel_predict(ref_pic(c, l, r)->hp, c->mb_w * 16, c->mb_h * 16, x, y, mv[0], mv[1], w, h, out);
}

static void luma_pred(const coder_t *c, int r, int x, int y, int w, int h, const int16_t *mv, uint16_t *out) {
  luma_pred_l(c, 0, r, x, y, w, h, mv, out);
}

/* enc/frame.rs:685-691 */
static int ref_bits(const coder_t *c, int r) {
  if (c->nref <= 1) return 0;
  if (c->nref == 2) return 1;
  int v = r + 1, lg = 0;
  while (v >>= 1) lg++;
  return 2 * lg + 1;
}

/* enc/frame.rs:944-947: the middle vector of the co-located macroblock of reference 0 (of a B
 * slice: of its list 1 picture, :1557), if inter */
static int col_mv(const coder_t *c, int mbx, int mby, int16_t *mv) {
  const ia_mbinfo *m = c->plist[c->b_slice ? 1 : 0].mbs;
  if (!m) return 0;
  const ia_mbinfo *col = &m[mby * c->mb_w + mbx];
  if (col->kind != K_INTER) return 0;
  mv[0] = col->mv[5][0];
  mv[1] = col->mv[5][1];
  return 1;
}

/* Index (ia_motion_search numbering) of the partition at (x, y, w, h) in 4x4 units. */
static int fs_index(int x, int y, int w, int h) {
  if (w == 4 && h == 4) return 0;
  if (w == 4 && h == 2) return 1 + y / 2;
  if (w == 2 && h == 4) return 3 + x / 2;
  int b8 = (y / 2) * 2 + x / 2;
  if (w == 2 && h == 2) return 5 + b8;
  if (w == 2 && h == 1) return 9 + 2 * b8 + (y & 1);
  if (w == 1 && h == 2) return 17 + 2 * b8 + (x & 1);
  return 25 + 4 * b8 + (y & 1) * 2 + (x & 1);
}

/* enc/frame.rs:949-1019 in list l, with IA_SEARCH_FULL's exhaustive winner as a last candidate (P
 * slices) */
static int motion_search_l(const coder_t *c, int l, int r, int mbx, int mby, int bx, int by, int bw, int bh,
                           const int16_t *mvp, const int16_t (*extra)[2], int nextra, int16_t *out) {
  const int lam = c->lambda;
  const int x0 = mbx * 16 + bx * 4, y0 = mby * 16 + by * 4, w = bw * 4, h = bh * 4;
  const int lo_x = maxi(-(x0 + 24) * 4, -2048), hi_x = mini((c->mb_w * 16 - x0 + 8) * 4, 2047);
  const int lo_y = maxi(-(y0 + 24) * 4, -512), hi_y = mini((c->mb_h * 16 - y0 + 8) * 4, 511);
#define CLAMPV(o, vx, vy) do { (o)[0] = (int16_t)clampi((vx), lo_x, hi_x); (o)[1] = (int16_t)clampi((vy), lo_y, hi_y); } while (0)
#define INTV(o) do { (o)[0] = (int16_t)(((o)[0] + 2) & ~3); (o)[1] = (int16_t)(((o)[1] + 2) & ~3); } while (0)
  int16_t cands[16][2];
  int nc = 0;
  CLAMPV(cands[nc], mvp[0], mvp[1]); INTV(cands[nc]); nc++;
  cands[nc][0] = 0; cands[nc][1] = 0; nc++;
  static const int NB[3][2] = {{-1, 0}, {0, -1}, {4, -1}};
  for (int k = 0; k < 3; k++) {
    int rr;
    int16_t mv[2];
    if (nb_motion(c, l, mbx, mby, ZMV, NOREF, 0, NB[k][0], NB[k][1], &rr, mv) && rr == r) {
      CLAMPV(cands[nc], mv[0], mv[1]); INTV(cands[nc]); nc++;
    }
  }
  for (int k = 0; k < nextra; k++) { CLAMPV(cands[nc], extra[k][0], extra[k][1]); INTV(cands[nc]); nc++; }
  if (c->search_mode == IA_SEARCH_FULL && c->fs && !c->b_slice) {
    const ia_fsresult *f = &c->fs[((size_t)r * c->mb_w * c->mb_h + (size_t)mby * c->mb_w + mbx) * IA_NPART + fs_index(bx, by, bw, bh)];
    CLAMPV(cands[nc], f->mv[0], f->mv[1]); INTV(cands[nc]); nc++;
  }
  int16_t best[2] = {0, 0};
  int best_cost = 0x7fffffff;
#define COST_INT(mv) (sad_int(c, l, r, x0, y0, w, h, (mv)) + lam * mvd_bits((mv), mvp))
  for (int k = 0; k < nc; k++) {
    int cost = COST_INT(cands[k]);
    if (cost < best_cost) { best[0] = cands[k][0]; best[1] = cands[k][1]; best_cost = cost; }
  }
  /* the realtime decisions skip the 16-sample step (enc/frame.rs:973) */
  static const int STEPS[3] = {16, 8, 4};
  for (int si = c->fast ? 1 : 0; si < 3; si++) {
    int step = STEPS[si];
    int iters = step == 4 ? 32 : 4;
    for (int it = 0; it < iters; it++) {
      int moved = 0;
      const int D[4][2] = {{-step, 0}, {step, 0}, {0, -step}, {0, step}};
      for (int d = 0; d < 4; d++) {
        int16_t cc[2];
        CLAMPV(cc, best[0] + D[d][0], best[1] + D[d][1]);
        int cost = COST_INT(cc);
        if (cost < best_cost) { best[0] = cc[0]; best[1] = cc[1]; best_cost = cost; moved = 1; }
      }
      if (!moved) break;
    }
  }
  /* realtime: stop at an integer vector that already matches (enc/frame.rs:990-992) */
  if (c->fast && best_cost < w * h / 8 + lam) {
    out[0] = best[0];
    out[1] = best[1];
    return best_cost - lam * mvd_bits(best, mvp);
  }
  /* sub-sample refinement by SATD: 8 half samples around, then 4 quarter samples (:994-1017);
     realtime coding keeps to the half samples (:1006) */
  uint16_t pred[256];
#define COST_SUB(mv) (luma_pred_l(c, l, r, x0, y0, w, h, (mv), pred), satd_block(c, 0, x0, y0, w, h, pred) + lam * mvd_bits((mv), mvp))
  best_cost = COST_SUB(best);
  static const int SQ[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
  static const int DM[4][2] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
  for (int pass = 0; pass < (c->fast ? 1 : 2); pass++) {
    int step = pass == 0 ? 2 : 1, np = pass == 0 ? 8 : 4;
    const int (*pts)[2] = pass == 0 ? SQ : DM;
    int16_t centre[2] = {best[0], best[1]};
    for (int k = 0; k < np; k++) {
      int16_t cc[2];
      CLAMPV(cc, centre[0] + pts[k][0] * step, centre[1] + pts[k][1] * step);
      int cost = COST_SUB(cc);
      if (cost < best_cost) { best[0] = cc[0]; best[1] = cc[1]; best_cost = cost; }
    }
  }
  out[0] = best[0];
  out[1] = best[1];
  return best_cost - lam * mvd_bits(best, mvp);
#undef CLAMPV
#undef INTV
#undef COST_INT
#undef COST_SUB
}

static int motion_search(const coder_t *c, int r, int mbx, int mby, int bx, int by, int bw, int bh,
                         const int16_t *mvp, const int16_t (*extra)[2], int nextra, int16_t *out) {
  return motion_search_l(c, 0, r, mbx, mby, bx, by, bw, bh, mvp, extra, nextra, out);
}

typedef struct {
  int16_t mvs[16][2];
  int8_t refs[16];
  uint32_t done;
} substate_t;

/* enc/frame.rs:911-940 */
static int sub_search(const coder_t *c, int mbx, int mby, int r, int k, const int16_t *extra, substate_t *st,
                      int *cost_out) {
  const int lam = c->lambda, ox = (k & 1) * 2, oy = (k >> 1) * 2;
  int best_s = -1, best_cost = 0;
  substate_t best_t;
  for (int s = 0; s < (c->sub8 ? 4 : 1); s++) {
    substate_t t = *st;
    int cost = lam * (s == 0 ? 0 : 2);
    for (int q = 0; q < NSUBS[s]; q++) {
      int x = ox + SUBS[s][q][0], y = oy + SUBS[s][q][1], w = SUBS[s][q][2], h = SUBS[s][q][3];
      int16_t mvp[2], m[2];
      pred_mv(c, 0, mbx, mby, (const int16_t(*)[2])t.mvs, t.refs, t.done, x, y, w, 0, r, mvp);
      int cc = motion_search(c, r, mbx, mby, x, y, w, h, mvp, (const int16_t(*)[2])extra, 1, m);
      cost += cc + lam * mvd_bits(m, mvp);
      for (int by = y; by < y + h; by++)
        for (int bx = x; bx < x + w; bx++) {
          t.mvs[by * 4 + bx][0] = m[0];
          t.mvs[by * 4 + bx][1] = m[1];
          t.refs[by * 4 + bx] = (int8_t)r;
          t.done |= 1u << (by * 4 + bx);
        }
    }
    if (best_s < 0 || cost < best_cost) { best_s = s; best_cost = cost; best_t = t; }
  }
  *st = best_t;
  *cost_out = best_cost;
  return best_s;
}

typedef struct {
  int part;
  int16_t mvs[16][2];
  uint8_t refs[4], sub[4];
  int cost;
} inter_t;

/* enc/frame.rs:852-906 */
static void best_inter_ref(const coder_t *c, int mbx, int mby, int r, inter_t *out) {
  const int lam = c->lambda, rb = lam * ref_bits(c, r);
  uint8_t refs[4] = {(uint8_t)r, (uint8_t)r, (uint8_t)r, (uint8_t)r};
  int16_t mvp[2], mv[2], col[1][2];
  mv_pred(c, 0, mbx, mby, r, mvp);
  int ncol = col_mv(c, mbx, mby, col[0]);
  int cost = motion_search(c, r, mbx, mby, 0, 0, 4, 4, mvp, (const int16_t(*)[2])col, ncol, mv);
  inter_t best;
  memset(&best, 0, sizeof best);
  best.part = 0;
  for (int i = 0; i < 16; i++) { best.mvs[i][0] = mv[0]; best.mvs[i][1] = mv[1]; }
  best.cost = cost + lam * (mvd_bits(mv, mvp) + 1) + rb;
  if (c->fast) {  /* realtime: 16x16 only (enc/frame.rs:861-863) */
    *out = best;
    memcpy(out->refs, refs, 4);
    return;
  }
  substate_t st;
  memset(&st, 0, sizeof st);
  uint8_t sub[4] = {0, 0, 0, 0};
  int cost8 = lam * 9;
  for (int k = 0; k < 4; k++) {
    int cc;
    sub[k] = (uint8_t)sub_search(c, mbx, mby, r, k, mv, &st, &cc);
    cost8 += cc + rb;
    if (cost8 >= best.cost) break;
  }
#define AT8(b) (st.mvs[((b) / 2) * 8 + ((b) % 2) * 2])
  if (cost8 < best.cost) {
    inter_t b8;
    memset(&b8, 0, sizeof b8);
    rect_t rs[16];
    int n = rects(3, sub, rs);
    for (int k = 0; k < n; k++) { b8.mvs[k][0] = st.mvs[rs[k].y * 4 + rs[k].x][0]; b8.mvs[k][1] = st.mvs[rs[k].y * 4 + rs[k].x][1]; }
    b8.part = 3;
    memcpy(b8.sub, sub, 4);
    b8.cost = cost8;
    best = b8;
    for (int part = 1; part < 3; part++) {
      int16_t mvs[16][2] = {{0}};
      int pc = lam * 3;
      static const uint8_t Z4[4] = {0, 0, 0, 0};
      for (int k = 0; k < 2; k++) {
        int16_t ps[16][2], m[2], cands[2][2];
        part_mvps(c, mbx, mby, part, Z4, (const int16_t(*)[2])mvs, refs, ps);
        int b0 = part == 1 ? 2 * k : k, b1 = part == 1 ? 2 * k + 1 : k + 2;
        cands[0][0] = AT8(b0)[0]; cands[0][1] = AT8(b0)[1];
        cands[1][0] = AT8(b1)[0]; cands[1][1] = AT8(b1)[1];
        const uint8_t *pr = PARTS[part][k];
        int cc = motion_search(c, r, mbx, mby, pr[0], pr[1], pr[2], pr[3], ps[k], (const int16_t(*)[2])cands, 2, m);
        mvs[k][0] = m[0];
        mvs[k][1] = m[1];
        pc += cc + lam * mvd_bits(m, ps[k]) + rb;
        if (pc >= best.cost) break;
      }
      if (pc < best.cost) {
        memset(&best, 0, sizeof best);
        best.part = part;
        memcpy(best.mvs, mvs, sizeof mvs);
        best.cost = pc;
      }
    }
  }
#undef AT8
  *out = best;
  memcpy(out->refs, refs, 4);
}

/* enc/frame.rs:839-848 */
static void best_inter(const coder_t *c, int mbx, int mby, inter_t *out) {
  int have = 0;
  for (int r = 0; r < c->nref; r++) {
    inter_t t;
    best_inter_ref(c, mbx, mby, r, &t);
    if (!have || t.cost < out->cost) {
      *out = t;
      for (int k = 0; k < 4; k++) out->refs[k] = (uint8_t)r;
      have = 1;
    }
  }
}

/* enc/frame.rs:792-810 */
static void i16_neighbours(const coder_t *c, int mbx, int mby, int32_t *top, int32_t *left, int32_t *tl) {
  const int x0 = mbx * 16, y0 = mby * 16, ls = ls_(c);
  const uint16_t *p = c->pic[0];
  memset(top, 0, 16 * sizeof *top);
  memset(left, 0, 16 * sizeof *left);
  if (up_(c, mby))
    for (int i = 0; i < 16; i++) top[i] = p[(size_t)(y0 - 1) * ls + x0 + i];
  if (mbx > 0)
    for (int j = 0; j < 16; j++) left[j] = p[(size_t)(y0 + j) * ls + x0 - 1];
  *tl = (mbx > 0 && up_(c, mby)) ? p[(size_t)(y0 - 1) * ls + x0 - 1] : 0;
}

/* enc/frame.rs:812-834 */
static int best_i16(const coder_t *c, int mbx, int mby, int *cost_out) {
  int ta = up_(c, mby), la = mbx > 0;
  int32_t top[16], left[16], tl;
  i16_neighbours(c, mbx, mby, top, left, &tl);
  int best_m = 2, best_c = 0x7fffffff;
  for (int mode = 0; mode < 4; mode++) {
    int ok = mode == 0 ? ta : mode == 1 ? la : mode == 2 ? 1 : (ta && la);
    if (!ok) continue;
    uint16_t pred[256];
    or_predict_16x16(mode, top, left, tl, ta, la, pred, c->bd);
    int cost = satd_block(c, 0, mbx * 16, mby * 16, 16, 16, pred) + c->lambda * (mode == 2 ? 0 : 1);
    if (cost < best_c) { best_m = mode; best_c = cost; }
  }
  *cost_out = best_c;
  return best_m;
}

/* enc/frame.rs:1025-1046 */
static void quant_block(const int32_t *src, const int32_t *pred, int qp, int intra, int skip_dc, int32_t *lv,
                        int32_t *co, int32_t *dc) {
  int32_t d[16], c[16];
  for (int k = 0; k < 16; k++) d[k] = src[k] - pred[k];
  or_fdct4(d, c);
  int qbits = 15 + qp / 6, f = (1 << qbits) / (intra ? 3 : 6), m = qp % 6;
  memset(lv, 0, 16 * sizeof *lv);
  memset(co, 0, 16 * sizeof *co);
  for (int k = 0; k < 16; k++) {
    int p = OR_ZIGZAG4[k];
    if (skip_dc && p == 0) continue;
    int l = or_quant(c[p], OR_MF[m][class4(p)], f, qbits);
    lv[k] = l;
    co[p] = shl(l * OR_V[m][class4(p)], qp / 6);
  }
  *dc = c[0];
}

static int count_nz(const int32_t *v, int n) {
  int k = 0;
  for (int i = 0; i < n; i++) k += v[i] != 0;
  return k;
}

/* Raster (bx, by) of the n-th 4x4 luma block in decoding order (enc/frame.rs:108-110). */
static void block_pos(int n, int *bx, int *by) {
  *bx = ((n >> 2) & 1) * 2 + (n & 1);
  *by = ((n >> 3) & 1) * 2 + ((n >> 1) & 1);
}

static int zorder(int bx, int by) { return ((by >> 1) << 3) | ((bx >> 1) << 2) | ((by & 1) << 1) | (bx & 1); }

/* enc/frame.rs:1050-1138 */
static int code_i4(coder_t *c, int mbx, int mby, mbcode_t *code) {
  mbcode_new(code, K_I4);
  const int ta = up_(c, mby), la = mbx > 0, tla = mbx > 0 && ta, tra = ta && mbx + 1 < c->mb_w, ls = ls_(c);
  int cost = 0;
  for (int n = 0; n < 16; n++) {
    int bx, by;
    block_pos(n, &bx, &by);
    const int x = mbx * 16 + bx * 4, y = mby * 16 + by * 4;
    int have_left = bx > 0 || la, have_top = by > 0 || ta, have_tl, have_tr;
    if (bx > 0 && by > 0) have_tl = 1;
    else if (by > 0) have_tl = la;
    else if (bx > 0) have_tl = ta;
    else have_tl = tla;
    if (by == 0) have_tr = bx < 3 ? ta : tra;
    else if (bx == 3) have_tr = 0;
    else have_tr = zorder(bx + 1, by - 1) < zorder(bx, by);
    const uint16_t *p = c->pic[0];
    edges_t e;
    memset(&e, 0, sizeof e);
    e.have_top = have_top; e.have_left = have_left; e.have_topleft = have_tl;
    if (have_top) {
      for (int i = 0; i < 4; i++) e.top[i] = p[(size_t)(y - 1) * ls + x + i];
      for (int i = 4; i < 8; i++) e.top[i] = have_tr ? p[(size_t)(y - 1) * ls + x + i] : e.top[3];
    }
    if (have_left)
      for (int j = 0; j < 4; j++) e.left[j] = p[(size_t)(y + j) * ls + x - 1];
    if (have_tl) e.topleft = p[(size_t)(y - 1) * ls + x - 1];
    int predicted = pred_i4_mode(c, code->modes, mbx, mby, bx, by);
    int best_m = 2, best_c = 0x7fffffff;
    uint16_t best_p[16] = {0};
    for (int mode = 0; mode < 9; mode++) {
      int ok;
      if (mode == 0 || mode == 3 || mode == 7) ok = have_top;
      else if (mode == 1 || mode == 8) ok = have_left;
      else if (mode == 2) ok = 1;
      else ok = have_top && have_left && have_tl;
      if (!ok) continue;
      uint16_t pred[16];
      predict_nxn(4, mode, &e, pred, c->bd);
      int cc = satd_block(c, 0, x, y, 4, 4, pred) + c->lambda * (mode == predicted ? 1 : 4);
      if (cc < best_c) { best_m = mode; best_c = cc; memcpy(best_p, pred, sizeof pred); }
    }
    cost += best_c;
    code->modes[by * 4 + bx] = (uint8_t)best_m;
    int32_t s[16], pr[16], lv[16], co[16], dc;
    src_block(c, 0, x, y, 4, s);
    for (int k = 0; k < 16; k++) pr[k] = best_p[k];
    quant_block(s, pr, qpp_(c), 1, 0, lv, co, &dc);
    uint16_t *pp = c->pic[0];
    for (int r = 0; r < 4; r++)
      for (int k = 0; k < 4; k++) pp[(size_t)(y + r) * ls + x + k] = best_p[r * 4 + k];
    int nz = count_nz(lv, 16);
    if (nz) or_idct4_add(co, pp + (size_t)y * ls + x, ls, maxv_(c));
    code->nnz[by * 4 + bx] = (uint8_t)nz;
    memcpy(code->luma[by * 4 + bx], lv, sizeof lv);
  }
  for (int i8 = 0; i8 < 4; i8++) {
    int bx = (i8 & 1) * 2, by = (i8 >> 1) * 2, any = 0;
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) any |= code->nnz[(by + dy) * 4 + bx + dx] > 0;
    if (any) code->cbp |= 1 << i8;
  }
  return cost;
}

/* enc/frame.rs:1140-1197 */
static void code_i16(coder_t *c, int mbx, int mby, int mode, mbcode_t *code) {
  mbcode_new(code, K_I16);
  code->i16_mode = mode;
  int32_t top[16], left[16], tl;
  i16_neighbours(c, mbx, mby, top, left, &tl);
  uint16_t pred[256];
  or_predict_16x16(mode, top, left, tl, up_(c, mby), mbx > 0, pred, c->bd);
  const int x0 = mbx * 16, y0 = mby * 16, ls = ls_(c), qp = qpp_(c);
  int32_t dcs[16], cos[16][16];
  for (int b = 0; b < 16; b++) {
    int bx = b % 4, by = b / 4;
    int32_t s[16], pr[16], lv[16];
    src_block(c, 0, x0 + bx * 4, y0 + by * 4, 4, s);
    for (int r = 0; r < 4; r++)
      for (int k = 0; k < 4; k++) pr[r * 4 + k] = pred[(by * 4 + r) * 16 + bx * 4 + k];
    quant_block(s, pr, qp, 1, 1, lv, cos[b], &dcs[b]);
    memcpy(code->luma[b], lv, sizeof lv);
  }
  int32_t h[16], dl[16] = {0};
  or_hadamard4(dcs, h);
  int qbits = 16 + qp / 6, f = (1 << qbits) / 3, m = qp % 6;
  for (int k = 0; k < 16; k++) {
    int p = OR_ZIGZAG4[k];
    code->dc[k] = or_quant(h[p] >> 1, OR_MF[m][0], f, qbits);
    dl[p] = code->dc[k];
  }
  int ac_coded = 0;
  for (int b = 0; b < 16; b++) ac_coded |= count_nz(code->luma[b], 16) > 0;
  if (!ac_coded) {
    memset(code->luma, 0, sizeof code->luma);
    memset(cos, 0, sizeof cos);
  }
  code->cbp = ac_coded ? 15 : 0;
  int32_t ff[16];
  or_hadamard4(dl, ff);
  int lsd = 16 * OR_V[m][0];
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) c->pic[0][(size_t)(y0 + y) * ls + x0 + x] = pred[y * 16 + x];
  for (int b = 0; b < 16; b++) {
    int32_t co[16];
    memcpy(co, cos[b], sizeof co);
    co[0] = qp >= 36 ? shl(ff[b] * lsd, qp / 6 - 6) : (ff[b] * lsd + (1 << (5 - qp / 6))) >> (6 - qp / 6);
    if (count_nz(co, 16)) {
      int bx = b % 4, by = b / 4;
      or_idct4_add(co, c->pic[0] + (size_t)(y0 + by * 4) * ls + x0 + bx * 4, ls, maxv_(c));
    }
    code->nnz[b] = (uint8_t)count_nz(code->luma[b], 16);
  }
}

/* enc/frame.rs:1856-1918 */
static void code_chroma(coder_t *c, mbcode_t *code, int mbx, int mby, const uint16_t (*pred)[64], int intra) {
  const int x0 = mbx * 8, y0 = mby * 8, cs = cs_(c), qp = c->qpc + 6 * (c->bd - 8), m = qp % 6;
  int32_t cos[2][4][16];
  int any_dc = 0, any_ac = 0;
  for (int ci = 0; ci < 2; ci++) {
    int32_t dcs[4];
    for (int blk = 0; blk < 4; blk++) {
      int bx = blk & 1, by = blk >> 1;
      int32_t s[16], pr[16];
      src_block(c, 1 + ci, x0 + bx * 4, y0 + by * 4, 4, s);
      for (int r = 0; r < 4; r++)
        for (int k = 0; k < 4; k++) pr[r * 4 + k] = pred[ci][(by * 4 + r) * 8 + bx * 4 + k];
      quant_block(s, pr, qp, intra, 1, code->cac[ci][blk], cos[ci][blk], &dcs[blk]);
      any_ac |= count_nz(code->cac[ci][blk], 16) > 0;
    }
    int32_t f[4] = {dcs[0] + dcs[1] + dcs[2] + dcs[3], dcs[0] - dcs[1] + dcs[2] - dcs[3],
                    dcs[0] + dcs[1] - dcs[2] - dcs[3], dcs[0] - dcs[1] - dcs[2] + dcs[3]};
    int qbits = 16 + qp / 6, fr = (1 << qbits) / (intra ? 3 : 6);
    for (int k = 0; k < 4; k++) code->cdc[ci][k] = or_quant(f[k], OR_MF[m][0], fr, qbits);
    any_dc |= count_nz(code->cdc[ci], 4) > 0;
  }
  int ac_score = 0;
  for (int ci = 0; ci < 2; ci++)
    for (int blk = 0; blk < 4; blk++) ac_score += or_decimate_score(code->cac[ci][blk]);
  if (!intra && ac_score < 7) any_ac = 0;
  if (!any_ac) {
    memset(code->cac, 0, sizeof code->cac);
    memset(cos, 0, sizeof cos);
  }
  int cbpc = any_ac ? 2 : (any_dc ? 1 : 0);
  code->cbp |= cbpc << 4;
  int lsd = 16 * OR_V[m][0];
  for (int ci = 0; ci < 2; ci++) {
    const int32_t *l = code->cdc[ci];
    int32_t f[4] = {l[0] + l[1] + l[2] + l[3], l[0] - l[1] + l[2] - l[3], l[0] + l[1] - l[2] - l[3], l[0] - l[1] - l[2] + l[3]};
    uint16_t *plane = c->pic[1 + ci];
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) plane[(size_t)(y0 + y) * cs + x0 + x] = pred[ci][y * 8 + x];
    for (int blk = 0; blk < 4; blk++) {
      int32_t co[16];
      memcpy(co, cos[ci][blk], sizeof co);
      co[0] = cbpc > 0 ? shl(f[blk] * lsd, qp / 6) >> 5 : 0;
      code->cnnz[ci][blk] = (uint8_t)count_nz(code->cac[ci][blk], 16);
      if (count_nz(co, 16)) {
        int bx = blk & 1, by = blk >> 1;
        or_idct4_add(co, plane + (size_t)(y0 + by * 4) * cs + x0 + bx * 4, cs, maxv_(c));
      }
    }
  }
}

/* enc/frame.rs:1254-1304 */
static void code_luma8(coder_t *c, mbcode_t *code, int mbx, int mby, const uint16_t *pred) {
  const int x0 = mbx * 16, y0 = mby * 16, ls = ls_(c);
  int32_t cos[4][64];
  int score[4];
  for (int i8 = 0; i8 < 4; i8++) {
    int bx = (i8 & 1) * 8, by = (i8 >> 1) * 8;
    int32_t s[64], pr[64];
    src_block(c, 0, x0 + bx, y0 + by, 8, s);
    for (int k = 0; k < 64; k++) pr[k] = pred[(by + k / 8) * 16 + bx + k % 8];
    or_quant8(s, pr, qpp_(c), 0, code->luma8[i8], cos[i8]);
    score[i8] = or_decimate_score8(code->luma8[i8]);
  }
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) c->pic[0][(size_t)(y0 + y) * ls + x0 + x] = pred[y * 16 + x];
  int total = score[0] + score[1] + score[2] + score[3];
  for (int i8 = 0; i8 < 4; i8++) {
    if (score[i8] < 4 || total < 6) {
      memset(code->luma8[i8], 0, sizeof code->luma8[i8]);
      continue;
    }
    int n = count_nz(code->luma8[i8], 64);
    if (n > 0) {
      or_idct8_add(cos[i8], c->pic[0] + (size_t)(y0 + (i8 >> 1) * 8) * ls + x0 + (i8 & 1) * 8, ls, maxv_(c));
      code->cbp |= 1 << i8;
    }
    for (int j = 0; j < 4; j++) code->nnz[((i8 >> 1) * 2 + (j >> 1)) * 4 + (i8 & 1) * 2 + (j & 1)] = (uint8_t)n;
  }
  code->t8x8 = (code->cbp & 15) != 0;
}

/* enc/frame.rs:1783-1805 */
static void inter_pred(const coder_t *c, int mbx, int mby, int part, const uint8_t *sub, const int16_t (*mvs)[2],
                       const uint8_t *refs, uint16_t *pred, uint16_t (*cpred)[64]) {
  const int x0 = mbx * 16, y0 = mby * 16, cs = cs_(c);
  rect_t rs[16];
  int n = rects(part, sub, rs);
  uint16_t tmp[256];
  for (int k = 0; k < n; k++) {
    const int bx = rs[k].x, by = rs[k].y, w = rs[k].w, h = rs[k].h;
    int r = refs[ref_of(part, k, bx, by)];
    luma_pred(c, r, x0 + bx * 4, y0 + by * 4, w * 4, h * 4, mvs[k], tmp);
    for (int rr = 0; rr < h * 4; rr++)
      for (int q = 0; q < w * 4; q++) pred[(by * 4 + rr) * 16 + bx * 4 + q] = tmp[rr * w * 4 + q];
    for (int ci = 0; ci < 2; ci++) {
      or_chroma_mc(c->plist[r].planes[1 + ci], cs, c->mb_h * 8, mbx * 8 + bx * 2, mby * 8 + by * 2, mvs[k][0], mvs[k][1], w * 2, h * 2, tmp);
      for (int rr = 0; rr < h * 2; rr++)
        for (int q = 0; q < w * 2; q++) cpred[ci][(by * 2 + rr) * 8 + bx * 2 + q] = tmp[rr * w * 2 + q];
    }
  }
}

/* enc/frame.rs:1217-1252 */
static void code_inter_residual(coder_t *c, mbcode_t *code, int mbx, int mby, const uint16_t *pred,
                                const uint16_t (*cpred)[64], int t8) {
  const int x0 = mbx * 16, y0 = mby * 16, ls = ls_(c);
  if (t8) {
    code_luma8(c, code, mbx, mby, pred);
  } else {
    int32_t cos[16][16];
    int score[4] = {0, 0, 0, 0};
    for (int b = 0; b < 16; b++) {
      int bx = b % 4, by = b / 4;
      int32_t s[16], pr[16], dc;
      src_block(c, 0, x0 + bx * 4, y0 + by * 4, 4, s);
      for (int r = 0; r < 4; r++)
        for (int k = 0; k < 4; k++) pr[r * 4 + k] = pred[(by * 4 + r) * 16 + bx * 4 + k];
      quant_block(s, pr, qpp_(c), 0, 0, code->luma[b], cos[b], &dc);
      score[(by / 2) * 2 + bx / 2] += or_decimate_score(code->luma[b]);
    }
    int total = score[0] + score[1] + score[2] + score[3];
    for (int i8 = 0; i8 < 4; i8++)
      if (score[i8] < 4 || total < 6)
        for (int dy = 0; dy < 2; dy++)
          for (int dx = 0; dx < 2; dx++) {
            int b = ((i8 >> 1) * 2 + dy) * 4 + (i8 & 1) * 2 + dx;
            memset(code->luma[b], 0, sizeof code->luma[b]);
            memset(cos[b], 0, sizeof cos[b]);
          }
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) c->pic[0][(size_t)(y0 + y) * ls + x0 + x] = pred[y * 16 + x];
    for (int b = 0; b < 16; b++) {
      code->nnz[b] = (uint8_t)count_nz(code->luma[b], 16);
      if (code->nnz[b] > 0) {
        int bx = b % 4, by = b / 4;
        or_idct4_add(cos[b], c->pic[0] + (size_t)(y0 + by * 4) * ls + x0 + bx * 4, ls, maxv_(c));
        code->cbp |= 1 << ((by / 2) * 2 + bx / 2);
      }
    }
  }
  code_chroma(c, code, mbx, mby, cpred, 0);
}

/* enc/frame.rs:1201-1214 */
static void code_inter(coder_t *c, int mbx, int mby, int part, const int16_t (*mvs)[2], const uint8_t *refs,
                       const uint8_t *sub, int t8, mbcode_t *code) {
  mbcode_new(code, K_INTER);
  code->part = part;
  memcpy(code->mvs, mvs, sizeof code->mvs);
  memcpy(code->refs, refs, 4);
  memcpy(code->sub, sub, 4);
  int16_t mvps[16][2];
  memset(mvps, 0, sizeof mvps);
  part_mvps(c, mbx, mby, part, sub, mvs, refs, mvps);
  for (int n = 0; n < 16; n++)
    for (int k = 0; k < 2; k++) code->mvds[n][k] = (int)mvs[n][k] - mvps[n][k];
  uint16_t pred[256], cpred[2][64];
  inter_pred(c, mbx, mby, part, sub, mvs, refs, pred, cpred);
  code_inter_residual(c, code, mbx, mby, pred, (const uint16_t(*)[64])cpred, t8);
}

/* enc/frame.rs:1307-1388 */
static void code_i8(coder_t *c, int mbx, int mby, mbcode_t *code) {
  mbcode_new(code, K_I8);
  code->t8x8 = 1;
  const int ta = up_(c, mby), la = mbx > 0, tla = mbx > 0 && ta, tra = ta && mbx + 1 < c->mb_w, ls = ls_(c);
  for (int i8 = 0; i8 < 4; i8++) {
    const int bx = i8 & 1, by = i8 >> 1, x = mbx * 16 + bx * 8, y = mby * 16 + by * 8;
    int have_left = bx > 0 || la, have_top = by > 0 || ta, have_tl, have_tr;
    if (bx > 0 && by > 0) have_tl = 1;
    else if (by > 0) have_tl = la;
    else if (bx > 0) have_tl = ta;
    else have_tl = tla;
    have_tr = i8 == 0 ? ta : i8 == 1 ? tra : i8 == 2 ? 1 : 0;
    const uint16_t *p = c->pic[0];
    edges_t e;
    memset(&e, 0, sizeof e);
    e.have_top = have_top; e.have_left = have_left; e.have_topleft = have_tl;
    if (have_top) {
      for (int i = 0; i < 8; i++) e.top[i] = p[(size_t)(y - 1) * ls + x + i];
      for (int i = 8; i < 16; i++) e.top[i] = have_tr ? p[(size_t)(y - 1) * ls + x + i] : e.top[7];
    }
    if (have_left)
      for (int j = 0; j < 8; j++) e.left[j] = p[(size_t)(y + j) * ls + x - 1];
    if (have_tl) e.topleft = p[(size_t)(y - 1) * ls + x - 1];
    filter_8x8_edges(&e);
    int predicted = pred_i4_mode(c, code->modes, mbx, mby, bx * 2, by * 2);
    int best_m = 2, best_c = 0x7fffffff;
    uint16_t best_p[64] = {0};
    for (int mode = 0; mode < 9; mode++) {
      int ok;
      if (mode == 0 || mode == 3 || mode == 7) ok = have_top;
      else if (mode == 1 || mode == 8) ok = have_left;
      else if (mode == 2) ok = 1;
      else ok = have_top && have_left && have_tl;
      if (!ok) continue;
      uint16_t pred[64];
      predict_nxn(8, mode, &e, pred, c->bd);
      int cc = satd_block(c, 0, x, y, 8, 8, pred) + c->lambda * (mode == predicted ? 1 : 4);
      if (cc < best_c) { best_m = mode; best_c = cc; memcpy(best_p, pred, sizeof pred); }
    }
    int32_t s[64], pr[64], lv[64], co[64];
    src_block(c, 0, x, y, 8, s);
    for (int k = 0; k < 64; k++) pr[k] = best_p[k];
    or_quant8(s, pr, qpp_(c), 1, lv, co);
    uint16_t *pp = c->pic[0];
    for (int r = 0; r < 8; r++)
      for (int k = 0; k < 8; k++) pp[(size_t)(y + r) * ls + x + k] = best_p[r * 8 + k];
    int n = count_nz(lv, 64);
    if (n > 0) {
      or_idct8_add(co, pp + (size_t)y * ls + x, ls, maxv_(c));
      code->cbp |= 1 << i8;
    }
    for (int j = 0; j < 4; j++) {
      int b = (by * 2 + (j >> 1)) * 4 + bx * 2 + (j & 1);
      code->modes[b] = (uint8_t)best_m;
      code->nnz[b] = (uint8_t)n;
    }
    memcpy(code->luma8[i8], lv, sizeof lv);
  }
}

/* enc/frame.rs:1807-1852 */
static void code_intra_chroma(coder_t *c, mbcode_t *code, int mbx, int mby) {
  const int la = mbx > 0, ta = up_(c, mby), x0 = mbx * 8, y0 = mby * 8, cs = cs_(c);
  int32_t top[2][8], left[2][16], tl[2];
  for (int ci = 0; ci < 2; ci++) {
    const uint16_t *p = c->pic[1 + ci];
    memset(top[ci], 0, sizeof top[ci]);
    memset(left[ci], 0, sizeof left[ci]);
    tl[ci] = 0;
    if (ta)
      for (int i = 0; i < 8; i++) top[ci][i] = p[(size_t)(y0 - 1) * cs + x0 + i];
    if (la)
      for (int j = 0; j < 8; j++) left[ci][j] = p[(size_t)(y0 + j) * cs + x0 - 1];
    if (la && ta) tl[ci] = p[(size_t)(y0 - 1) * cs + x0 - 1];
  }
  int best_m = 0, best_c = 0x7fffffff;
  uint16_t best_p[2][64];
  memset(best_p, 0, sizeof best_p);
  for (int mode = 0; mode < 4; mode++) {
    int ok = mode == 0 ? 1 : mode == 1 ? la : mode == 2 ? ta : (la && ta);
    if (!ok) continue;
    uint16_t preds[2][64];
    int cost = c->lambda * (mode == 0 ? 1 : 3);
    for (int ci = 0; ci < 2; ci++) {
      uint16_t out[128];
      or_predict_chroma(mode, top[ci], left[ci], tl[ci], ta, la, out, c->bd, 8);
      memcpy(preds[ci], out, 64 * sizeof *out);
      cost += satd_block(c, 1 + ci, x0, y0, 8, 8, preds[ci]);
    }
    if (cost < best_c) { best_m = mode; best_c = cost; memcpy(best_p, preds, sizeof preds); }
  }
  code->chroma_mode = best_m;
  code_chroma(c, code, mbx, mby, (const uint16_t(*)[64])best_p, 1);
}

/* enc/frame.rs:1401-1430 */
static void save_mb(const coder_t *c, int mbx, int mby, uint16_t *rec) {
  const int ls = ls_(c), cs = cs_(c);
  for (int r = 0; r < 16; r++)
    for (int k = 0; k < 16; k++) rec[r * 16 + k] = c->pic[0][(size_t)(mby * 16 + r) * ls + mbx * 16 + k];
  for (int ci = 0; ci < 2; ci++)
    for (int r = 0; r < 8; r++)
      for (int k = 0; k < 8; k++) rec[256 + ci * 64 + r * 8 + k] = c->pic[1 + ci][(size_t)(mby * 8 + r) * cs + mbx * 8 + k];
}

static void restore_mb(coder_t *c, int mbx, int mby, const uint16_t *rec) {
  const int ls = ls_(c), cs = cs_(c);
  for (int r = 0; r < 16; r++)
    for (int k = 0; k < 16; k++) c->pic[0][(size_t)(mby * 16 + r) * ls + mbx * 16 + k] = rec[r * 16 + k];
  for (int ci = 0; ci < 2; ci++)
    for (int r = 0; r < 8; r++)
      for (int k = 0; k < 8; k++) c->pic[1 + ci][(size_t)(mby * 8 + r) * cs + mbx * 8 + k] = rec[256 + ci * 64 + r * 8 + k];
}

/* enc/frame.rs:1433-1449 */
static int64_t ssd_mb(const coder_t *c, int mbx, int mby) {
  uint16_t rec[384];
  save_mb(c, mbx, mby, rec);
  const int ls = ls_(c), cs = cs_(c);
  int64_t ssd = 0;
  for (int i = 0; i < 384; i++) {
    int s;
    if (i < 256) s = c->src[0][(size_t)(mby * 16 + i / 16) * ls + mbx * 16 + i % 16];
    else {
      int ci = (i - 256) / 64, j = (i - 256) % 64;
      s = c->src[1 + ci][(size_t)(mby * 8 + j / 8) * cs + mbx * 8 + j % 8];
    }
    int64_t d = (int64_t)s - rec[i];
    ssd += d * d;
  }
  return ssd;
}

static void mvd_grid(const rect_t *rs, int n, const int32_t (*mvds)[2], uint8_t (*out)[2]) {
  memset(out, 0, 16 * 2);
  for (int k = 0; k < n; k++)
    for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
      for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++)
        for (int q = 0; q < 2; q++) out[by * 4 + bx][q] = (uint8_t)mini(absi(mvds[k][q]), 64);
}

static int b_mb_type(const mbcode_t *code);

/* enc/frame.rs:2028-2279 */
static void mb_bins(const coder_t *c, bins_t *b, const mbcode_t *code, int mbx, int mby) {
  const mbstate_t *left = nb_state(c, mbx - 1, mby), *top = nb_state(c, mbx, mby - 1);
  if (c->p_slice) {
    decision(b, (c->b_slice ? 24 : 11) + (left && !left->skip) + (top && !top->skip), code->skip);
    if (code->skip) return;
  }
  const int cbpl = code->cbp & 15, cbpc = code->cbp >> 4, prefix_ok = c->p_slice;
  const int base = c->b_slice ? 32 : c->p_slice ? 17 : 3;
  const int b_ctx = 27 + (left && !left->bdirect) + (top && !top->bdirect);
  if (code->kind == K_INTER && c->b_slice) {
    /* B_Direct_16x16 0, B_L0_16x16 100, B_L1_16x16 101, the others 11 and four or five more bins
       (Table 9-37) */
    const int t = b_mb_type(code);
    decision(b, b_ctx, t != 0);
    if (t == 1 || t == 2) {
      decision(b, 30, 0);
      decision(b, 32, t == 2);
    } else if (t != 0) {
      decision(b, 30, 1);
      const int bits = t >= 3 && t <= 10 ? t - 3 : t == 11 ? 14 : t == 22 ? 15 : (t + 4) >> 1;
      decision(b, 31, (bits >> 3) & 1);
      for (int sh = 2; sh >= 0; sh--) decision(b, 32, (bits >> sh) & 1);
      if (t >= 12 && t <= 21) decision(b, 32, (t + 4) & 1);
    }
    if (t == 22)
      /* sub_mb_type: B_Direct_8x8 0, B_L0_8x8 100, B_L1_8x8 101, B_Bi_8x8 11000 */
      for (int k = 0; k < 4; k++) {
        const int d = code->bdir[k];
        decision(b, 36, d != 0);
        if (d == 1 || d == 2) {
          decision(b, 37, 0);
          decision(b, 39, d == 2);
        } else if (d == 3) {
          decision(b, 37, 1);
          decision(b, 38, 0);
          decision(b, 39, 0);
          decision(b, 39, 0);
        }
      }
  } else if (code->kind == K_INTER) {
    static const int B1[4] = {0, 1, 1, 0}, B2[4] = {0, 1, 0, 1};
    decision(b, 14, 0);
    decision(b, 15, B1[code->part]);
    decision(b, B1[code->part] ? 17 : 16, B2[code->part]);
  } else {
    int first;
    if (c->b_slice) {
      /* 111101: an intra type follows, with the suffix contexts */
      decision(b, b_ctx, 1);
      decision(b, 30, 1);
      decision(b, 31, 1);
      decision(b, 32, 1);
      decision(b, 32, 0);
      decision(b, 32, 1);
      first = 32;
    } else if (prefix_ok) {
      decision(b, 14, 1);
      first = 17;
    } else {
      first = 3 + (left && !(left->kind == K_I4 || left->kind == K_I8)) + (top && !(top->kind == K_I4 || top->kind == K_I8));
    }
    if (code->kind == K_I4 || code->kind == K_I8) {
      decision(b, first, 0);
      if (c->transform8) decision(b, 399 + (left && left->t8x8) + (top && top->t8x8), code->kind == K_I8);
    } else {
      decision(b, first, 1);
      terminate(b, 0);
      int b1 = prefix_ok ? base + 1 : base + 3, b2 = prefix_ok ? base + 2 : base + 4, b3 = prefix_ok ? base + 3 : base + 6;
      decision(b, b1, cbpl != 0);
      decision(b, b2, cbpc != 0);
      if (cbpc != 0) decision(b, prefix_ok ? base + 2 : base + 5, cbpc == 2);
      decision(b, b3, (code->i16_mode >> 1) == 1);
      decision(b, prefix_ok ? base + 3 : base + 7, (code->i16_mode & 1) == 1);
    }
  }
  if (code->kind == K_I4 || code->kind == K_I8) {
    const int t8 = code->kind == K_I8;
    uint8_t so_far[16];
    memset(so_far, 2, 16);
    for (int n = 0; n < (t8 ? 4 : 16); n++) {
      int bx, by;
      if (t8) { bx = (n & 1) * 2; by = (n >> 1) * 2; }
      else block_pos(n, &bx, &by);
      int mode = code->modes[by * 4 + bx];
      int pred = pred_i4_mode(c, so_far, mbx, mby, bx, by);
      so_far[by * 4 + bx] = (uint8_t)mode;
      if (t8) {
        so_far[by * 4 + bx + 1] = (uint8_t)mode;
        so_far[(by + 1) * 4 + bx] = (uint8_t)mode;
        so_far[(by + 1) * 4 + bx + 1] = (uint8_t)mode;
      }
      decision(b, 68, mode == pred);
      if (mode != pred) {
        int rem = mode < pred ? mode : mode - 1;
        for (int k = 0; k < 3; k++) decision(b, 69, (rem >> k) & 1);
      }
    }
  }
  if (code->kind != K_INTER) {
    int m = code->chroma_mode;
    int ca = left && !left->inter && left->chroma_mode != 0, cb = top && !top->inter && top->chroma_mode != 0;
    decision(b, 64 + ca + cb, m > 0);
    if (m > 0) {
      decision(b, 67, m > 1);
      if (m > 1) decision(b, 67, m > 2);
    }
  } else {
    if (!c->b_slice && code->part == 3) {
      for (int k = 0; k < 4; k++) {
        int s = code->sub[k];
        if (s == 0) decision(b, 21, 1);
        else if (s == 1) { decision(b, 21, 0); decision(b, 22, 0); }
        else if (s == 2) { decision(b, 21, 0); decision(b, 22, 1); decision(b, 23, 1); }
        else { decision(b, 21, 0); decision(b, 22, 1); decision(b, 23, 0); }
      }
    }
    if (!c->b_slice && c->nref > 1) {
      uint8_t grid[4] = {0, 0, 0, 0};
      rect_t parts[4];
      int np;
      if (code->part == 3) {
        np = 4;
        for (int k = 0; k < 4; k++) parts[k] = (rect_t){(k & 1) * 2, (k >> 1) * 2, 2, 2};
      } else {
        np = NPARTS[code->part];
        for (int k = 0; k < np; k++) parts[k] = (rect_t){PARTS[code->part][k][0], PARTS[code->part][k][1], PARTS[code->part][k][2], PARTS[code->part][k][3]};
      }
      for (int k = 0; k < np; k++) {
        int bx = parts[k].x, by = parts[k].y, w = parts[k].w, h = parts[k].h;
        int a = bx > 0 ? grid[(by / 2) * 2 + (bx - 1) / 2] > 0 : (left && left->refs[(by / 2) * 2 + 1] > 0);
        int bb = by > 0 ? grid[((by - 1) / 2) * 2 + bx / 2] > 0 : (top && top->refs[2 + bx / 2] > 0);
        int r = code->refs[k], cbase = 54 + a + 2 * bb;
        for (int i = 0; i <= r; i++) decision(b, i == 0 ? cbase : (i == 1 ? 58 : 59), i < r);
        for (int y8 = by / 2; y8 <= (by + h - 1) / 2; y8++)
          for (int x8 = bx / 2; x8 <= (bx + w - 1) / 2; x8++) grid[y8 * 2 + x8] = code->refs[k];
      }
    }
    /* list 0's differences for the partitions using it, then list 1's */
    static const uint8_t Z4[4] = {0, 0, 0, 0};
    rect_t rs[16];
    int n = rects(code->part, c->b_slice ? Z4 : code->sub, rs);
    for (int l = 0; l < (c->b_slice ? 2 : 1); l++) {
      const int32_t (*mvds)[2] = c->b_slice ? code->bmvd[l] : code->mvds;
      uint8_t grid[16][2];
      mvd_grid(rs, n, mvds, grid);
      for (int k = 0; k < n; k++) {
        if (c->b_slice && !((code->bdir[k] >> l) & 1)) continue;
        int bx = rs[k].x, by = rs[k].y;
        for (int comp = 0; comp < 2; comp++) {
          int a = bx > 0 ? grid[by * 4 + bx - 1][comp] : (left ? (l ? left->mvd1 : left->mvd)[by * 4 + 3][comp] : 0);
          int bb = by > 0 ? grid[(by - 1) * 4 + bx][comp] : (top ? (l ? top->mvd1 : top->mvd)[12 + bx][comp] : 0);
          mvd_bins(b, comp, mvds[k][comp], (uint32_t)(a + bb));
        }
      }
    }
  }
  if (code->kind != K_I16) {
    int luma = 0;
    for (int b8 = 0; b8 < 4; b8++) {
      int cond_a = (b8 & 1) ? ((luma >> (b8 - 1)) & 1) == 0 : (left && ((left->cbp >> (b8 + 1)) & 1) == 0);
      int cond_b = b8 >= 2 ? ((luma >> (b8 - 2)) & 1) == 0 : (top && ((top->cbp >> (b8 + 2)) & 1) == 0);
      int bit = (cbpl >> b8) & 1;
      decision(b, 73 + cond_a + 2 * cond_b, bit);
      luma |= bit << b8;
    }
    int ca = left && (left->cbp >> 4) != 0, cb = top && (top->cbp >> 4) != 0;
    decision(b, 77 + ca + 2 * cb, cbpc != 0);
    if (cbpc != 0) {
      ca = left && (left->cbp >> 4) == 2;
      cb = top && (top->cbp >> 4) == 2;
      decision(b, 81 + ca + 2 * cb, cbpc == 2);
    }
  }
  if (code->kind == K_INTER && c->transform8 && cbpl != 0 &&
      (code->part != 3 || (code->sub[0] | code->sub[1] | code->sub[2] | code->sub[3]) == 0))
    decision(b, 399 + (left && left->t8x8) + (top && top->t8x8), code->t8x8);
  if (code->cbp != 0 || code->kind == K_I16) {
    decision(b, 60, 0);
    const int un = code->kind != K_INTER;
#define FLAG_NB(avail, v) ((avail) ? ((v) > 0) : un)
    if (code->kind == K_I16) {
      int da = left ? (left->cbf_dc & 1) != 0 : un, db = top ? (top->cbf_dc & 1) != 0 : un;
      write_block(b, 0, 85 + da + 2 * db, code->dc, 16);
    }
    const int cat = code->kind == K_I16 ? 1 : 2, cbase = code->kind == K_I16 ? 85 + 4 : 85 + 8;
    for (int i8 = 0; i8 < 4; i8++)
      if (code->t8x8 && (cbpl & (1 << i8))) write_block(b, 5, -1, code->luma8[i8], 64);
    for (int n = 0; n < 16; n++) {
      int bx, by;
      block_pos(n, &bx, &by);
      if (code->t8x8 || !(cbpl & (1 << (n / 4)))) continue;
      int a = bx > 0 ? code->nnz[by * 4 + bx - 1] > 0 : FLAG_NB(left != NULL, left ? left->nnz[by * 4 + 3] : 0);
      int bb = by > 0 ? code->nnz[(by - 1) * 4 + bx] > 0 : FLAG_NB(top != NULL, top ? top->nnz[12 + bx] : 0);
      const int32_t *blk = code->luma[by * 4 + bx];
      if (cat == 1) write_block(b, cat, cbase + a + 2 * bb, blk + 1, 15);
      else write_block(b, cat, cbase + a + 2 * bb, blk, 16);
    }
    if (cbpc > 0)
      for (int comp = 0; comp < 2; comp++) {
        int da = left ? (left->cbf_dc & (2 << comp)) != 0 : un, db = top ? (top->cbf_dc & (2 << comp)) != 0 : un;
        write_block(b, 3, 85 + 12 + da + 2 * db, code->cdc[comp], 4);
      }
    if (cbpc == 2)
      for (int comp = 0; comp < 2; comp++)
        for (int blk = 0; blk < 4; blk++) {
          int bx = blk & 1, by = blk >> 1;
          const uint8_t *cur = code->cnnz[comp];
          int a = bx > 0 ? cur[by * 2] > 0 : FLAG_NB(left != NULL, left ? left->cnnz[comp][by * 2 + 1] : 0);
          int bb = by > 0 ? cur[bx] > 0 : FLAG_NB(top != NULL, top ? top->cnnz[comp][2 + bx] : 0);
          write_block(b, 4, 85 + 16 + a + 2 * bb, code->cac[comp][blk] + 1, 15);
        }
#undef FLAG_NB
  }
}

/* enc/frame.rs:1453-1457 */
static double rd_cost(const coder_t *c, const mbcode_t *code, int mbx, int mby, double lam2) {
  uint8_t st[1024];
  memcpy(st, c->cabac, sizeof st);
  bins_t b = {st, 0};
  mb_bins(c, &b, code, mbx, mby);
  double bits = (double)b.bits / 256.0;
  return (double)ssd_mb(c, mbx, mby) + lam2 * bits;
}

typedef struct {
  double j;
  mbcode_t code;
  uint16_t rec[384];
  int have;
} keep_t;

static void keep(coder_t *c, keep_t *k, double j, const mbcode_t *code, int mbx, int mby) {
  if (!k->have || j < k->j) {
    k->have = 1;
    k->j = j;
    k->code = *code;
    save_mb(c, mbx, mby, k->rec);
  }
}

/* enc/frame.rs:1462-1513 */
static void decide_rd(coder_t *c, int mbx, int mby, int i16, const inter_t *inter, mbcode_t *out) {
  const double lam2 = 1.6 * pow(2.0, ((double)qpp_(c) - 12.0) / 3.0);
  static _Thread_local keep_t k;
  k.have = 0;
  static _Thread_local mbcode_t code;
  if (inter) {
    int16_t sm[2];
    skip_mv(c, mbx, mby, sm);
    mbcode_new(&code, K_INTER);
    code.skip = 1;
    for (int i = 0; i < 16; i++) { code.mvs[i][0] = sm[0]; code.mvs[i][1] = sm[1]; }
    {
      uint16_t pred[256], cpred[2][64];
      static const uint8_t Z4[4] = {0, 0, 0, 0};
      inter_pred(c, mbx, mby, 0, Z4, (const int16_t(*)[2])code.mvs, Z4, pred, cpred);
      uint16_t rec[384];
      memcpy(rec, pred, sizeof pred);
      memcpy(rec + 256, cpred, sizeof cpred);
      restore_mb(c, mbx, mby, rec);
    }
    keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
    code_inter(c, mbx, mby, inter->part, (const int16_t(*)[2])inter->mvs, inter->refs, inter->sub, 0, &code);
    if (!(code.cbp == 0 && inter->part == 0 && inter->mvs[0][0] == sm[0] && inter->mvs[0][1] == sm[1] && inter->refs[0] == 0))
      keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
    if (c->transform8 && (inter->sub[0] | inter->sub[1] | inter->sub[2] | inter->sub[3]) == 0) {
      code_inter(c, mbx, mby, inter->part, (const int16_t(*)[2])inter->mvs, inter->refs, inter->sub, 1, &code);
      if (code.t8x8) keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
    }
  }
  code_i16(c, mbx, mby, i16, &code);
  code_intra_chroma(c, &code, mbx, mby);
  keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  code_i4(c, mbx, mby, &code);
  code_intra_chroma(c, &code, mbx, mby);
  keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  if (c->transform8) {
    code_i8(c, mbx, mby, &code);
    code_intra_chroma(c, &code, mbx, mby);
    keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  }
  restore_mb(c, mbx, mby, k.rec);
  *out = k.code;
  c->st->rd_cost = k.j;
}

/* ---- B slices (enc/frame.rs:1515-1780) ---------------------------------------------------------- */

/* The 8x8 block (raster) holding raster 4x4 block b. */
static int b8_of(int b) { return (b / 8) * 2 + (b % 4) / 2; }

/* The smallest non-negative of two reference indices, or the larger (-1) when neither is (8.4.1.2.2
 * MinPositive, enc/frame.rs:1724). */
static int min_positive(int x, int y) { return x >= 0 && y >= 0 ? mini(x, y) : maxi(x, y); }

/* Spatial direct motion of a macroblock (enc/frame.rs:1717-1745, 8.4.1.2.2): per list the neighbours'
 * smallest reference index and the 16x16 vector predictor, zeroed in the 8x8 blocks whose co-located
 * corner block of the list 1 picture uses its reference 0 and moves by at most one quarter sample;
 * both lists with zero vectors when no neighbour uses either. */
static void direct_mvs(const coder_t *c, int mbx, int mby, int16_t (*d)[16][2], uint8_t *use) {
  int refs[2];
  for (int l = 0; l < 2; l++) {
    int ra = -1, rb = -1, rc = -1, r;
    int16_t m[2];
    if (nb_motion(c, l, mbx, mby, ZMV, NOREF, 0, -1, 0, &r, m)) ra = r;
    if (nb_motion(c, l, mbx, mby, ZMV, NOREF, 0, 0, -1, &r, m)) rb = r;
    if (nb_motion(c, l, mbx, mby, ZMV, NOREF, 0, 4, -1, &r, m)) rc = r;
    else if (nb_motion(c, l, mbx, mby, ZMV, NOREF, 0, -1, -1, &r, m)) rc = r;
    refs[l] = min_positive(ra, min_positive(rb, rc));
  }
  memset(d, 0, 2 * 16 * 2 * sizeof(int16_t));
  if (refs[0] < 0 && refs[1] < 0) {
    use[0] = use[1] = 1;
    return;
  }
  for (int l = 0; l < 2; l++)
    if (refs[l] >= 0) {
      int16_t p[2];
      mv_pred(c, l, mbx, mby, 0, p);
      for (int b = 0; b < 16; b++) { d[l][b][0] = p[0]; d[l][b][1] = p[1]; }
    }
  const ia_mbinfo *m1 = c->plist[1].mbs;
  const ia_mbinfo *col = m1 ? &m1[mby * c->mb_w + mbx] : NULL;
  if (col && col->kind == K_INTER)
    for (int b = 0; b < 16; b++) {
      /* the corner 4x4 block of each 8x8 block stands for it (direct_8x8_inference_flag) */
      const int cx = ((b % 4) >> 1) * 3, cy = ((b / 4) >> 1) * 3;
      const int16_t *mv = col->mv[cy * 4 + cx];
      if (col->ref_idx[(cy >> 1) * 2 + (cx >> 1)] == 0 && absi(mv[0]) <= 1 && absi(mv[1]) <= 1)
        for (int l = 0; l < 2; l++)
          if (refs[l] == 0) d[l][b][0] = d[l][b][1] = 0;
    }
  use[0] = refs[0] >= 0;
  use[1] = refs[1] >= 0;
}

/* The prediction of a B macroblock 8x8 block by 8x8 block from the lists used there, averaged when
 * both are (enc/frame.rs:1749-1780). */
static void b_pred(const coder_t *c, int mbx, int mby, const int16_t (*mv)[16][2], const uint8_t (*used)[2],
                   uint16_t *pred, uint16_t (*cpred)[64]) {
  const int x0 = mbx * 16, y0 = mby * 16, cs = cs_(c);
  for (int i8 = 0; i8 < 4; i8++) {
    const int bx = (i8 & 1) * 8, by = (i8 >> 1) * 8;
    uint16_t lp[2][64], cp[2][2][16];
    memset(lp, 0, sizeof lp);
    memset(cp, 0, sizeof cp);
    for (int l = 0; l < 2; l++) {
      if (!used[i8][l]) continue;
      const int16_t *v = mv[l][(by / 4) * 4 + bx / 4];
      luma_pred_l(c, l, 0, x0 + bx, y0 + by, 8, 8, v, lp[l]);
      for (int ci = 0; ci < 2; ci++)
        or_chroma_mc(c->plist[l].planes[1 + ci], cs, c->mb_h * 8, (x0 + bx) / 2, (y0 + by) / 2, v[0], v[1], 4, 4, cp[l][ci]);
    }
#define MIX(a, b) (used[i8][0] && used[i8][1] ? (uint16_t)(((a) + (b) + 1) >> 1) : used[i8][0] ? (a) : (b))
    for (int k = 0; k < 64; k++) pred[(by + k / 8) * 16 + bx + k % 8] = MIX(lp[0][k], lp[1][k]);
    for (int ci = 0; ci < 2; ci++)
      for (int k = 0; k < 16; k++) cpred[ci][(by / 2 + k / 4) * 8 + bx / 2 + k % 4] = MIX(cp[0][ci][k], cp[1][ci][k]);
#undef MIX
  }
}

/* Each list's vector differences per partition of a B macroblock (enc/frame.rs:1688-1711): partitions
 * predicted in order from the earlier ones that use the list, a direct 8x8 block counting with its
 * derived motion. */
static void b_mvds(const coder_t *c, int mbx, int mby, int part, const uint8_t *bdir, const int16_t (*bmv)[16][2],
                   const uint8_t (*buse)[2], int32_t (*out)[4][2]) {
  memset(out, 0, 2 * 4 * 2 * sizeof(int32_t));
  const int dir = (int[]){0, 1, 2, 0}[part];
  for (int l = 0; l < 2; l++) {
    int16_t cur[16][2];
    memset(cur, 0, sizeof cur);
    uint32_t done = 0;
    for (int k = 0; k < NPARTS[part]; k++) {
      const int x = PARTS[part][k][0], y = PARTS[part][k][1], w = PARTS[part][k][2], h = PARTS[part][k][3];
      if ((bdir[k] >> l) & 1) {
        int16_t mvp[2];
        pred_mv(c, l, mbx, mby, (const int16_t(*)[2])cur, NOREF, done, x, y, w, dir, 0, mvp);
        out[l][k][0] = (int)bmv[l][y * 4 + x][0] - mvp[0];
        out[l][k][1] = (int)bmv[l][y * 4 + x][1] - mvp[1];
      }
      for (int by = y; by < y + h; by++)
        for (int bx = x; bx < x + w; bx++) {
          const int b = by * 4 + bx;
          cur[b][0] = bmv[l][b][0];
          cur[b][1] = bmv[l][b][1];
          done |= 1u << b;
          if (!buse[b8_of(b)][l]) done |= 1u << (16 + b);
        }
    }
  }
}

/* One partition's best direction for best_b_parts (enc/frame.rs:1628-1656): list 0, list 1, both
 * averaged, and with `direct` B_Direct_8x8 from dpred; returns the direction, mv the lists' vectors. */
static int b_part_search(const coder_t *c, int mbx, int mby, int bx, int by, int w, int h, const int16_t (*extra0)[2],
                         int n0, const int16_t (*extra1)[2], int n1, int direct, const int16_t (*mvps)[2],
                         const uint16_t *dpred, int16_t (*mv)[2], int *cost_out) {
  const int lam = c->lambda, x0 = mbx * 16 + bx * 4, y0 = mby * 16 + by * 4, pw = w * 4, ph = h * 4;
  int cost[2];
  uint16_t preds[2][128], avg[128];
  for (int l = 0; l < 2; l++) {
    cost[l] = motion_search_l(c, l, 0, mbx, mby, bx, by, w, h, mvps[l], l ? extra1 : extra0, l ? n1 : n0, mv[l]);
    luma_pred_l(c, l, 0, x0, y0, pw, ph, mv[l], preds[l]);
  }
  for (int i = 0; i < pw * ph; i++) avg[i] = (uint16_t)((preds[0][i] + preds[1][i] + 1) >> 1);
  const int bi = satd_block(c, 0, x0, y0, pw, ph, avg) + lam * (mvd_bits(mv[0], mvps[0]) + mvd_bits(mv[1], mvps[1]) + 2);
  int bd, bc;
  if (bi < mini(cost[0], cost[1])) { bd = 3; bc = bi; }
  else if (cost[0] <= cost[1]) { bd = 1; bc = cost[0]; }
  else { bd = 2; bc = cost[1]; }
  if (direct) {
    /* sub_mb_type takes about one bin for direct and three for the others */
    bc += lam * 3;
    uint16_t sub[128];
    for (int i = 0; i < pw * ph; i++) sub[i] = dpred[(by * 4 + i / pw) * 16 + bx * 4 + i % pw];
    const int d = satd_block(c, 0, x0, y0, pw, ph, sub) + lam;
    if (d < bc) { bd = 0; bc = d; }
  }
  *cost_out = bc;
  return bd;
}

typedef struct {
  int part;
  uint8_t dirs[4];
  int16_t pmv[2][4][2];
} bparts_t;

/* The best B partitioning below 16x16 when it costs less than cost16 (enc/frame.rs:1621-1685): each
 * 8x8 block direct or from list 0, list 1 or both, then 16x8 and 8x16 (without direct); returns 0
 * when the 8x8 blocks' running cost reaches cost16. */
static int best_b_parts(const coder_t *c, int mbx, int mby, const int16_t (*mvs16)[2], const int16_t (*mvps)[2],
                        int cost16, const uint16_t *dpred, bparts_t *out) {
  const int lam = c->lambda;
  uint8_t dirs[4] = {0, 0, 0, 0};
  int16_t pmv[2][4][2];
  memset(pmv, 0, sizeof pmv);
  int total = lam * 6;
  for (int k = 0; k < 4; k++) {
    int16_t mv[2][2];
    int cc;
    dirs[k] = (uint8_t)b_part_search(c, mbx, mby, PARTS[3][k][0], PARTS[3][k][1], 2, 2, &mvs16[0], 1, &mvs16[1], 1, 1,
                                     mvps, dpred, mv, &cc);
    for (int l = 0; l < 2; l++) { pmv[l][k][0] = mv[l][0]; pmv[l][k][1] = mv[l][1]; }
    total += cc;
    if (total >= cost16) return 0;
  }
  out->part = 3;
  memcpy(out->dirs, dirs, 4);
  memcpy(out->pmv, pmv, sizeof pmv);
  int best_total = total;
  int16_t extra[2][5][2];
  for (int l = 0; l < 2; l++) {
    for (int k = 0; k < 4; k++) { extra[l][k][0] = pmv[l][k][0]; extra[l][k][1] = pmv[l][k][1]; }
    extra[l][4][0] = mvs16[l][0];
    extra[l][4][1] = mvs16[l][1];
  }
  for (int part = 1; part < 3; part++) {
    uint8_t d2[4] = {0, 0, 0, 0};
    int16_t m2[2][4][2];
    memset(m2, 0, sizeof m2);
    int t2 = lam * 7;
    for (int k = 0; k < 2; k++) {
      int16_t mv[2][2];
      int cc;
      const uint8_t *r = PARTS[part][k];
      d2[k] = (uint8_t)b_part_search(c, mbx, mby, r[0], r[1], r[2], r[3], (const int16_t(*)[2])extra[0], 5,
                                     (const int16_t(*)[2])extra[1], 5, 0, mvps, dpred, mv, &cc);
      for (int l = 0; l < 2; l++) { m2[l][k][0] = mv[l][0]; m2[l][k][1] = mv[l][1]; }
      t2 += cc;
      if (t2 >= best_total) break;
    }
    if (t2 < best_total) {
      out->part = part;
      memcpy(out->dirs, d2, 4);
      memcpy(out->pmv, m2, sizeof m2);
      best_total = t2;
    }
  }
  return 1;
}

/* The mb_type of a B macroblock (Table 7-14; enc/frame.rs:2343-2351). */
static int b_mb_type(const mbcode_t *code) {
  static const uint8_t PAIRS[9][2] = {{1, 1}, {2, 2}, {1, 2}, {2, 1}, {1, 3}, {2, 3}, {3, 1}, {3, 2}, {3, 3}};
  if (code->part == 0) return code->bdir[0];
  if (code->part == 3) return 22;
  int i = 0;
  while (i < 9 && !(PAIRS[i][0] == code->bdir[0] && PAIRS[i][1] == code->bdir[1])) i++;
  return 4 + 2 * i + (code->part == 2);
}

static void put_pred(coder_t *c, int mbx, int mby, const uint16_t *pred, const uint16_t (*cpred)[64]) {
  uint16_t rec[384];
  memcpy(rec, pred, 256 * sizeof *rec);
  memcpy(rec + 256, cpred, 128 * sizeof *rec);
  restore_mb(c, mbx, mby, rec);
}

/* A B macroblock by rate-distortion cost among B_Skip and B_Direct_16x16, the best of list 0, list 1
 * and both at 16x16 and the best smaller partitioning, Intra16x16 and Intra4x4 (enc/frame.rs:1518-1617). */
static void decide_b(coder_t *c, int mbx, int mby, mbcode_t *out) {
  const int lam = c->lambda, x0 = mbx * 16, y0 = mby * 16;
  const double lam2 = 1.6 * pow(2.0, ((double)qpp_(c) - 12.0) / 3.0);
  static _Thread_local keep_t k;
  k.have = 0;
  static _Thread_local mbcode_t code;
  const int nt8 = c->transform8 ? 2 : 1;
  int16_t direct[2][16][2];
  uint8_t duse[2], dused[4][2];
  direct_mvs(c, mbx, mby, direct, duse);
  for (int i = 0; i < 4; i++) { dused[i][0] = duse[0]; dused[i][1] = duse[1]; }
  uint16_t dpred[256], dcpred[2][64];
  b_pred(c, mbx, mby, (const int16_t(*)[16][2])direct, (const uint8_t(*)[2])dused, dpred, dcpred);
  mbcode_new(&code, K_INTER);
  code.skip = 1;
  memcpy(code.buse, dused, sizeof dused);
  memcpy(code.bmv, direct, sizeof direct);
  put_pred(c, mbx, mby, dpred, (const uint16_t(*)[64])dcpred);
  keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  for (int t8 = 0; t8 < nt8; t8++) {
    mbcode_new(&code, K_INTER);
    memcpy(code.buse, dused, sizeof dused);
    memcpy(code.bmv, direct, sizeof direct);
    code_inter_residual(c, &code, mbx, mby, dpred, (const uint16_t(*)[64])dcpred, t8);
    /* without residual, B_Direct_16x16 is B_Skip at a higher price */
    if (code.cbp != 0 && (!t8 || code.t8x8)) keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  }
  /* 16x16 from either list or both */
  int16_t mvps[2][2], mvs[2][2];
  int costs[2];
  uint16_t preds[2][256];
  mv_pred(c, 0, mbx, mby, 0, mvps[0]);
  mv_pred(c, 1, mbx, mby, 0, mvps[1]);
  const int tb = c->dist_tb, td = c->dist_td;
  int16_t colv[2];
  const int ncol = col_mv(c, mbx, mby, colv);
  for (int l = 0; l < 2; l++) {
    /* the co-located vector spans both lists' references, split by the picture's distances */
    int16_t extra[2][2];
    int ne = 0;
    if (ncol) {
      const int num = l == 0 ? tb : tb - td;
      extra[ne][0] = (int16_t)(colv[0] * num / td);
      extra[ne][1] = (int16_t)(colv[1] * num / td);
      ne++;
    }
    if (l == 1) {
      extra[ne][0] = (int16_t)(mvs[0][0] * (tb - td) / tb);
      extra[ne][1] = (int16_t)(mvs[0][1] * (tb - td) / tb);
      ne++;
    }
    const int cc = motion_search_l(c, l, 0, mbx, mby, 0, 0, 4, 4, mvps[l], (const int16_t(*)[2])extra, ne, mvs[l]);
    costs[l] = cc + lam * (mvd_bits(mvs[l], mvps[l]) + 3);
    luma_pred_l(c, l, 0, x0, y0, 16, 16, mvs[l], preds[l]);
  }
  uint16_t avg[256];
  for (int i = 0; i < 256; i++) avg[i] = (uint16_t)((preds[0][i] + preds[1][i] + 1) >> 1);
  const int bi = satd_block(c, 0, x0, y0, 16, 16, avg) + lam * (mvd_bits(mvs[0], mvps[0]) + mvd_bits(mvs[1], mvps[1]) + 5);
  const int dir = bi < mini(costs[0], costs[1]) ? 3 : costs[0] <= costs[1] ? 1 : 2;
  const int best16 = mini(bi, mini(costs[0], costs[1]));
  /* candidates: the 16x16 one, and the best smaller partitioning if it beats it */
  bparts_t cands[2];
  int ncand = 1;
  memset(cands, 0, sizeof cands);
  cands[0].part = 0;
  cands[0].dirs[0] = (uint8_t)dir;
  for (int l = 0; l < 2; l++)
    for (int q = 0; q < 4; q++) { cands[0].pmv[l][q][0] = mvs[l][0]; cands[0].pmv[l][q][1] = mvs[l][1]; }
  if (best_b_parts(c, mbx, mby, (const int16_t(*)[2])mvs, (const int16_t(*)[2])mvps, best16, dpred, &cands[1])) ncand = 2;
  c->st->inter_cost = best16;
  c->st->inter_part = cands[ncand - 1].part;
  for (int ci = 0; ci < ncand; ci++) {
    const bparts_t *cd = &cands[ci];
    uint8_t buse[4][2];
    int16_t bmv[2][16][2];
    int32_t bmvd[2][4][2];
    memset(buse, 0, sizeof buse);
    memset(bmv, 0, sizeof bmv);
    for (int q = 0; q < NPARTS[cd->part]; q++) {
      const uint8_t *r = PARTS[cd->part][q];
      for (int by = r[1]; by < r[1] + r[3]; by++)
        for (int bx = r[0]; bx < r[0] + r[2]; bx++) {
          const int b = by * 4 + bx, b8 = b8_of(b);
          for (int l = 0; l < 2; l++) {
            buse[b8][l] = cd->dirs[q] == 0 ? duse[l] : (cd->dirs[q] >> l) & 1;
            if (buse[b8][l]) {
              const int16_t *v = cd->dirs[q] == 0 ? direct[l][b] : cd->pmv[l][q];
              bmv[l][b][0] = v[0];
              bmv[l][b][1] = v[1];
            }
          }
        }
    }
    b_mvds(c, mbx, mby, cd->part, cd->dirs, (const int16_t(*)[16][2])bmv, (const uint8_t(*)[2])buse, bmvd);
    uint16_t pred[256], cpred[2][64];
    b_pred(c, mbx, mby, (const int16_t(*)[16][2])bmv, (const uint8_t(*)[2])buse, pred, cpred);
    for (int t8 = 0; t8 < nt8; t8++) {
      mbcode_new(&code, K_INTER);
      code.part = cd->part;
      memcpy(code.bdir, cd->dirs, 4);
      memcpy(code.buse, buse, sizeof buse);
      memcpy(code.bmv, bmv, sizeof bmv);
      memcpy(code.bmvd, bmvd, sizeof bmvd);
      code_inter_residual(c, &code, mbx, mby, pred, (const uint16_t(*)[64])cpred, t8);
      if (!t8 || code.t8x8) keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
    }
  }
  int i16c;
  const int i16 = best_i16(c, mbx, mby, &i16c);
  c->st->i16_cost = i16c;
  code_i16(c, mbx, mby, i16, &code);
  code_intra_chroma(c, &code, mbx, mby);
  keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  /* Intra8x8 is not tried in B pictures (enc/frame.rs:1611) */
  code_i4(c, mbx, mby, &code);
  code_intra_chroma(c, &code, mbx, mby);
  keep(c, &k, rd_cost(c, &code, mbx, mby, lam2), &code, mbx, mby);
  restore_mb(c, mbx, mby, k.rec);
  *out = k.code;
  c->st->rd_cost = k.j;
}

/* enc/frame.rs:715-790 */
static void code_mb(coder_t *c, int mbx, int mby, mbcode_t *out) {
  const int x0 = mbx * 16, y0 = mby * 16, lam = c->lambda;
  ia_mbstats *st = c->st;
  st->skip_sad = -1;
  st->i16_cost = -1;
  st->inter_cost = -1;
  st->inter_part = -1;
  st->rd_cost = 0.0;
  if (c->b_slice) {
    decide_b(c, mbx, mby, out);
    return;
  }
  if (c->p_slice) {
    int16_t smv[2];
    skip_mv(c, mbx, mby, smv);
    uint16_t pred[256];
    luma_pred(c, 0, x0, y0, 16, 16, smv, pred);
    const int ls = ls_(c);
    int sad = 0;
    for (int i = 0; i < 256; i++) sad += absi((int)c->src[0][(size_t)(y0 + i / 16) * ls + x0 + i % 16] - pred[i]);
    st->skip_sad = sad;
    double step = pow(2.0, (double)qpp_(c) / 6.0);
    if (c->fast && (double)sad < 256.0 * 0.05 * step) {
      /* realtime: P_Skip without a transform when luma and chroma are within a twentieth of a
         quantiser step of the skip prediction (enc/frame.rs:731-746) */
      int16_t mvs[16][2];
      for (int i = 0; i < 16; i++) { mvs[i][0] = smv[0]; mvs[i][1] = smv[1]; }
      static const uint8_t Z4[4] = {0, 0, 0, 0};
      uint16_t rec[384], (*cpred)[64] = (uint16_t(*)[64])(rec + 256);
      inter_pred(c, mbx, mby, 0, Z4, (const int16_t(*)[2])mvs, Z4, rec, cpred);
      const int cs = cs_(c);
      int csad = 0;
      for (int k = 0; k < 2; k++)
        for (int i = 0; i < 64; i++) csad += absi((int)c->src[1 + k][(size_t)(mby * 8 + i / 8) * cs + mbx * 8 + i % 8] - cpred[k][i]);
      if ((double)csad < 128.0 * 0.05 * step) {
        mbcode_new(out, K_INTER);
        out->skip = 1;
        memcpy(out->mvs, mvs, sizeof mvs);
        memcpy(rec, pred, sizeof pred);
        restore_mb(c, mbx, mby, rec);
        return;
      }
    }
    if ((double)sad < 256.0 * 0.625 * step) {
      int16_t mvs[16][2];
      for (int i = 0; i < 16; i++) { mvs[i][0] = smv[0]; mvs[i][1] = smv[1]; }
      static const uint8_t Z4[4] = {0, 0, 0, 0};
      code_inter(c, mbx, mby, 0, (const int16_t(*)[2])mvs, Z4, Z4, 0, out);
      if (out->cbp == 0) {
        out->skip = 1;
        return;
      }
    }
  }
  int i16c;
  int i16 = best_i16(c, mbx, mby, &i16c);
  st->i16_cost = i16c;
  int best_intra = i16c + lam * 4;
  inter_t inter;
  int have_inter = 0;
  if (c->p_slice) {
    best_inter(c, mbx, mby, &inter);
    have_inter = 1;
    st->inter_cost = inter.cost;
    st->inter_part = inter.part;
  }
  if (c->use_cabac) {
    decide_rd(c, mbx, mby, i16, have_inter ? &inter : NULL, out);
    return;
  }
  /* SATD decisions (CAVLC), enc/frame.rs:760-789 */
  int try_i4 = !have_inter || best_intra < inter.cost;
  /* realtime: flat intra macroblocks stay Intra16x16 (enc/frame.rs:763-765) */
  if (c->fast && !have_inter && (double)i16c < 256.0 * 0.8 * pow(2.0, (double)qpp_(c) / 6.0)) try_i4 = 0;
  static _Thread_local mbcode_t i4code;
  int have_i4 = 0;
  if (try_i4) {
    int cost = code_i4(c, mbx, mby, &i4code) + lam * 6;
    if (cost < best_intra) {
      best_intra = cost;
      have_i4 = 1;
    }
  }
  if (have_inter && inter.cost <= best_intra) {
    code_inter(c, mbx, mby, inter.part, (const int16_t(*)[2])inter.mvs, inter.refs, inter.sub, 0, out);
    int16_t sm[2];
    skip_mv(c, mbx, mby, sm);
    if (out->cbp == 0 && inter.part == 0 && inter.mvs[0][0] == sm[0] && inter.mvs[0][1] == sm[1] && inter.refs[0] == 0) out->skip = 1;
    return;
  }
  if (have_i4) *out = i4code;
  else code_i16(c, mbx, mby, i16, out);
  code_intra_chroma(c, out, mbx, mby);
}

/* enc/frame.rs:2281-2339 */
static void finish_mb(coder_t *c, const mbcode_t *code, int mbx, int mby) {
  const int idx = (mby - c->row0) * c->mb_w + mbx;
  const int inter = code->kind == K_INTER, bm = inter && c->b_slice;
  static const uint8_t Z4[4] = {0, 0, 0, 0};
  mbstate_t *s = &c->state[idx];
  memset(s, 0, sizeof *s);
  /* the lists each 8x8 block uses */
  uint8_t used[4][2];
  for (int b8 = 0; b8 < 4; b8++) {
    used[b8][0] = inter && (!bm || code->buse[b8][0]);
    used[b8][1] = bm && code->buse[b8][1];
  }
  rect_t rs[16];
  int n = rects(code->part, bm ? Z4 : code->sub, rs);
  if (inter && !code->skip) {
    if (bm) {
      mvd_grid(rs, n, code->bmvd[0], s->mvd);
      mvd_grid(rs, n, code->bmvd[1], s->mvd1);
    } else {
      mvd_grid(rs, n, code->mvds, s->mvd);
    }
  }
  if (bm) {
    for (int b = 0; b < 16; b++)
      for (int q = 0; q < 2; q++) {
        s->mvs[b][q] = used[b8_of(b)][0] ? code->bmv[0][b][q] : 0;
        s->mvs1[b][q] = used[b8_of(b)][1] ? code->bmv[1][b][q] : 0;
      }
  } else if (inter) {
    for (int k = 0; k < n; k++)
      for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
        for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++) {
          s->mvs[by * 4 + bx][0] = code->mvs[k][0];
          s->mvs[by * 4 + bx][1] = code->mvs[k][1];
        }
  }
  s->bdirect = (uint8_t)(bm && code->part == 0 && code->bdir[0] == 0);
  const int cbpc = code->cbp >> 4;
  s->cbf_dc = (uint8_t)((code->kind == K_I16 && count_nz(code->dc, 16) > 0) | ((cbpc > 0 && count_nz(code->cdc[0], 4) > 0) << 1) |
                        ((cbpc > 0 && count_nz(code->cdc[1], 4) > 0) << 2));
  int8_t r16[16];
  for (int k = 0; k < NPARTS[code->part]; k++) {
    const uint8_t *pr = PARTS[code->part][k];
    for (int by = pr[1]; by < pr[1] + pr[3]; by++)
      for (int bx = pr[0]; bx < pr[0] + pr[2]; bx++) r16[by * 4 + bx] = (int8_t)code->refs[k];
  }
  int8_t r8[4];
  for (int b8 = 0; b8 < 4; b8++) r8[b8] = r16[(b8 / 2) * 8 + (b8 % 2) * 2];
  s->i4 = code->kind == K_I4 || code->kind == K_I8;
  s->t8x8 = (uint8_t)code->t8x8;
  s->inter = (uint8_t)inter;
  memcpy(s->modes, code->modes, 16);
  memcpy(s->nnz, code->nnz, 16);
  memcpy(s->cnnz, code->cnnz, 8);
  for (int b8 = 0; b8 < 4; b8++) {
    s->refs[b8] = inter && !bm ? r8[b8] : used[b8][0] ? 0 : -1;
    s->refs1[b8] = used[b8][1] ? 0 : -1;
  }
  s->skip = (uint8_t)code->skip;
  s->kind = (uint8_t)code->kind;
  s->cbp = (uint8_t)code->cbp;
  s->chroma_mode = inter ? 0 : (uint8_t)code->chroma_mode;
  ia_mbinfo *info = &c->mbs[mby * c->mb_w + mbx];
  memset(info, 0, sizeof *info);
  info->slice = (uint16_t)c->slice_id;
  info->kind = (uint8_t)code->kind;
  info->qp = (int8_t)c->qp;
  memcpy(info->nnz, code->nnz, 16);
  info->cbp = (uint8_t)code->cbp;
  info->skip = (uint8_t)code->skip;
  info->t8x8 = (uint8_t)code->t8x8;
  memcpy(info->mv, s->mvs, sizeof info->mv);
  memcpy(info->mv1, s->mvs1, sizeof info->mv1);
  for (int b8 = 0; b8 < 4; b8++) {
    /* the reference index as the decoder's list gives it; the identity tells the loop filter which
       picture it is (B pictures use one picture per list) */
    const int r = bm ? 0 : r8[b8];
    info->ref_idx[b8] = used[b8][0] ? (int8_t)r : -1;
    info->ref_id[b8] = used[b8][0] ? (bm ? 1 : 1 + r) : -1;
    info->ref_idx1[b8] = used[b8][1] ? 0 : -1;
    info->ref_id1[b8] = used[b8][1] ? 2 : -1;
  }
  info->bdirect = s->bdirect;
  if (s->bdirect) info->direct8 = 15;
  else if (bm && code->part == 3)
    for (int k = 0; k < 4; k++)
      if (code->bdir[k] == 0) info->direct8 |= (uint8_t)(1 << k);
}

static void export_code(const mbcode_t *c, ia_mbcode *o) {
  memset(o, 0, sizeof *o);
  o->kind = (uint8_t)c->kind;
  o->skip = (uint8_t)c->skip;
  o->i16_mode = (uint8_t)c->i16_mode;
  o->chroma_mode = (uint8_t)c->chroma_mode;
  o->part = (uint8_t)c->part;
  o->t8x8 = (uint8_t)c->t8x8;
  o->cbp = (uint8_t)c->cbp;
  memcpy(o->modes, c->modes, 16);
  memcpy(o->refs, c->refs, 4);
  memcpy(o->sub, c->sub, 4);
  memcpy(o->nnz, c->nnz, 16);
  memcpy(o->cnnz, c->cnnz, 8);
  for (int i = 0; i < 16; i++)
    for (int k = 0; k < 2; k++) {
      o->mvs[i][k] = c->mvs[i][k];
      o->mvds[i][k] = (int16_t)c->mvds[i][k];
    }
  for (int b = 0; b < 16; b++)
    for (int k = 0; k < 16; k++) o->luma[b][k] = (int16_t)c->luma[b][k];
  for (int b = 0; b < 4; b++)
    for (int k = 0; k < 64; k++) o->luma8[b][k] = (int16_t)c->luma8[b][k];
  for (int k = 0; k < 16; k++) o->dc[k] = (int16_t)c->dc[k];
  for (int ci = 0; ci < 2; ci++)
    for (int k = 0; k < 4; k++) {
      o->cdc[ci][k] = (int16_t)c->cdc[ci][k];
      for (int q = 0; q < 16; q++) o->cac[ci][k][q] = (int16_t)c->cac[ci][k][q];
    }
  memcpy(o->bdir, c->bdir, 4);
  for (int b8 = 0; b8 < 4; b8++) o->buse[b8] = (uint8_t)(c->buse[b8][0] | c->buse[b8][1] << 1);
  for (int l = 0; l < 2; l++) {
    for (int b = 0; b < 16; b++) { o->bmv[l][b][0] = c->bmv[l][b][0]; o->bmv[l][b][1] = c->bmv[l][b][1]; }
    for (int k = 0; k < 4; k++) { o->bmvd[l][k][0] = (int16_t)c->bmvd[l][k][0]; o->bmvd[l][k][1] = (int16_t)c->bmvd[l][k][1]; }
  }
}

/* ---------------------------------------------------------------------------------------------
 * Full search (ia_motion_search's definition), brute force
 * ------------------------------------------------------------------------------------------- */

void or_fullsearch(const uint8_t *src_y, const uint8_t *const *ref_y, int nref, int mb_w, int mb_h, int range,
                   ia_fsresult *out) {
  const int pw = mb_w * 16, ph = mb_h * 16, sw = pw + 2 * OR_PAD, sh = ph + 2 * OR_PAD;
  uint8_t *g = malloc((size_t)sw * sh);
  for (int r = 0; r < nref; r++) {
    /* the edge-extended integer plane (enc/subpel.rs:74-77) */
    for (int y = 0; y < sh; y++)
      for (int x = 0; x < sw; x++)
        g[(size_t)y * sw + x] = ref_y[r][(size_t)clampi(y - OR_PAD, 0, ph - 1) * pw + clampi(x - OR_PAD, 0, pw - 1)];
    for (int mby = 0; mby < mb_h; mby++)
      for (int mbx = 0; mbx < mb_w; mbx++) {
        const int x0 = mbx * 16, y0 = mby * 16;
        const int dxlo = maxi(-range, -(x0 + 24)), dxhi = mini(range, pw - x0 + 8);
        const int dylo = maxi(-range, -(y0 + 24)), dyhi = mini(range, ph - y0 + 8);
        ia_fsresult *o = out + ((size_t)r * mb_w * mb_h + (size_t)mby * mb_w + mbx) * IA_NPART;
        for (int p = 0; p < IA_NPART; p++) { o[p].sad = 0xffffffffu; o[p].mv[0] = o[p].mv[1] = 0; }
        for (int dy = dylo; dy <= dyhi; dy++)
          for (int dx = dxlo; dx <= dxhi; dx++) {
            int s4[16];
            for (int b = 0; b < 16; b++) {
              int bx = (b & 3) * 4, by = (b >> 2) * 4, s = 0;
              for (int yy = 0; yy < 4; yy++)
                for (int xx = 0; xx < 4; xx++)
                  s += absi((int)src_y[(size_t)(y0 + by + yy) * pw + x0 + bx + xx] -
                            g[(size_t)(y0 + by + yy + dy + OR_PAD) * sw + x0 + bx + xx + dx + OR_PAD]);
              s4[b] = s;
            }
            uint32_t sad[IA_NPART];
#define B4(x, y) s4[(y) * 4 + (x)]
            sad[0] = 0;
            for (int b = 0; b < 16; b++) sad[0] += s4[b];
            for (int hh = 0; hh < 2; hh++) {
              sad[1 + hh] = 0;
              sad[3 + hh] = 0;
              for (int q = 0; q < 8; q++) {
                sad[1 + hh] += B4(q & 3, hh * 2 + (q >> 2));
                sad[3 + hh] += B4(hh * 2 + (q & 1), q >> 1);
              }
            }
            for (int b8 = 0; b8 < 4; b8++) {
              int ox = (b8 & 1) * 2, oy = (b8 >> 1) * 2;
              sad[5 + b8] = B4(ox, oy) + B4(ox + 1, oy) + B4(ox, oy + 1) + B4(ox + 1, oy + 1);
              sad[9 + 2 * b8] = B4(ox, oy) + B4(ox + 1, oy);
              sad[10 + 2 * b8] = B4(ox, oy + 1) + B4(ox + 1, oy + 1);
              sad[17 + 2 * b8] = B4(ox, oy) + B4(ox, oy + 1);
              sad[18 + 2 * b8] = B4(ox + 1, oy) + B4(ox + 1, oy + 1);
              sad[25 + 4 * b8] = B4(ox, oy);
              sad[26 + 4 * b8] = B4(ox + 1, oy);
              sad[27 + 4 * b8] = B4(ox, oy + 1);
              sad[28 + 4 * b8] = B4(ox + 1, oy + 1);
            }
#undef B4
            for (int p = 0; p < IA_NPART; p++)
              if (sad[p] < o[p].sad) { o[p].sad = sad[p]; o[p].mv[0] = (int16_t)(dx * 4); o[p].mv[1] = (int16_t)(dy * 4); }
          }
      }
  }
  free(g);
}

/* ---------------------------------------------------------------------------------------------
 * Picture driver: the slices of one picture (enc/mod.rs:539-660, analysis part)
 * ------------------------------------------------------------------------------------------- */

static uint16_t *widen(const uint8_t *p, size_t n) {
  uint16_t *o = malloc(n * sizeof *o);
  for (size_t i = 0; i < n; i++) o[i] = p[i];
  return o;
}

static int chroma_qp(int qp, int offset, int bd) {
  int qpi = clampi(qp + offset, -6 * (bd - 8), 51);
  return qpi < 30 ? qpi : OR_CHROMA_QP[qpi - 30];
}

/* One slice of a picture: the per-slice coder of enc/frame.rs:381-487 over its rows. */
typedef struct {
  const ia_config *cfg;
  uint16_t **src, **pic;
  refpic_t *refs;
  int np, qp, mb_w, mb_h, nslices, threads, first;
  int b_slice, tb, td;
  mbstate_t *state;
  const ia_fsresult *fs;
  ia_mbcode *codes;
  ia_mbstats *stats;
  ia_mbinfo *mbs;
} slice_job_t;

static void code_slice(const slice_job_t *j, int s, mbcode_t *code) {
  coder_t c;
  memset(&c, 0, sizeof c);
  c.src[0] = j->src[0]; c.src[1] = j->src[1]; c.src[2] = j->src[2];
  c.bd = 8;
  c.mb_w = j->mb_w;
  c.mb_h = j->mb_h;
  c.qp = j->qp;
  c.qpc = chroma_qp(j->qp, 0, 8);
  c.lambda = or_lambda(j->qp);
  c.row0 = s * j->mb_h / j->nslices;
  c.row1 = (s + 1) * j->mb_h / j->nslices;
  c.slice_id = s + 1;
  c.p_slice = j->np > 0;
  c.nref = j->b_slice ? 1 : j->np;
  c.b_slice = j->b_slice;
  c.dist_tb = j->tb;
  c.dist_td = j->td;
  c.plist = j->refs;
  c.sub8 = j->cfg->sub8x8 && j->qp <= 20;
  c.use_cabac = j->cfg->cabac;
  c.transform8 = j->cfg->cabac && j->cfg->transform8;
  c.pic[0] = j->pic[0]; c.pic[1] = j->pic[1]; c.pic[2] = j->pic[2];
  c.mbs = j->mbs;
  c.state = j->state + (size_t)c.row0 * j->mb_w;
  c.fs = j->fs;
  c.search_mode = j->cfg->search_mode;
  c.fast = j->cfg->fast != 0;
  cabac_init(c.cabac, !c.p_slice, j->qp);
  for (int mby = c.row0; mby < c.row1; mby++)
    for (int mbx = 0; mbx < j->mb_w; mbx++) {
      const int idx = mby * j->mb_w + mbx;
      ia_mbstats dummy;
      c.st = j->stats ? &j->stats[idx] : &dummy;
      code_mb(&c, mbx, mby, code);
      c.st->bits = 0;
      if (c.use_cabac) {
        bins_t live = {c.cabac, 0};
        mb_bins(&c, &live, code, mbx, mby);
        c.st->bits = (int32_t)live.bits;
      }
      finish_mb(&c, code, mbx, mby);
      if (j->codes) export_code(code, &j->codes[idx]);
    }
}

/* Worker t codes slices t, t + threads, ... (slices share nothing but the read-only references
 * and source; each writes its own rows of the picture, records and state). */
static void *slice_worker(void *arg) {
  const slice_job_t *j = arg;
  mbcode_t *code = malloc(sizeof *code);
  for (int s = j->first; s < j->nslices; s += j->threads) code_slice(j, s, code);
  free(code);
  return NULL;
}

static int analyse_impl(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp, int kind,
                        const uint8_t *const *ref_y, const uint8_t *const *ref_cb, const uint8_t *const *ref_cr,
                        const ia_mbinfo *const *ref_mbs, int nref, int tb, int td, const ia_fsresult *fs,
                        ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y, uint8_t *recon_cb, uint8_t *recon_cr,
                        ia_mbinfo *mbs, int threads) {
  const int mb_w = (cfg->width + 15) / 16, mb_h = (cfg->height + 15) / 16, pw = mb_w * 16, ph = mb_h * 16;
  const size_t ylen = (size_t)pw * ph, clen = ylen / 4;
  if (cfg->bit_depth != 8 || qp < 0 || qp > 51 || (kind == IA_PIC_P && (nref < 1 || nref > 16))) return -1;
  /* B pictures: one picture per list, CABAC only (enc/mod.rs:534-536), 0 < tb < td */
  if (kind == IA_PIC_B && (nref != 2 || !cfg->cabac || tb <= 0 || td <= tb)) return -1;
  if (kind != IA_PIC_IDR && kind != IA_PIC_P && kind != IA_PIC_B) return -1;
  entropy_init();
  const int nslices = clampi(cfg->slices, 1, mb_h);
  threads = clampi(threads, 1, nslices);
  uint16_t *src[3] = {widen(y, ylen), widen(cb, clen), widen(cr, clen)};
  uint16_t *pic[3] = {calloc(ylen, 2), calloc(clen, 2), calloc(clen, 2)};
  refpic_t refs[16];
  int np = kind == IA_PIC_IDR ? 0 : nref;
  for (int r = 0; r < np; r++) {
    uint16_t *ly = widen(ref_y[r], ylen);
    refs[r].hp = malloc((size_t)4 * (pw + 2 * OR_PAD) * (ph + 2 * OR_PAD) * 2);
    or_halfpel_build(ly, pw, ph, 8, refs[r].hp);
    free(ly);
    refs[r].planes[0] = NULL;
    refs[r].planes[1] = widen(ref_cb[r], clen);
    refs[r].planes[2] = widen(ref_cr[r], clen);
    refs[r].mbs = ref_mbs ? ref_mbs[r] : NULL;
  }
  mbstate_t *state = calloc((size_t)mb_w * mb_h, sizeof *state);
  slice_job_t jobs[64];
  pthread_t tids[64];
  if (threads > 64) threads = 64;
  int started = 0, rc = 0;
  for (int t = 0; t < threads; t++) {
    jobs[t] = (slice_job_t){cfg, src, pic, refs, np, qp, mb_w, mb_h, nslices, threads, t, kind == IA_PIC_B, tb, td,
                            state, fs, codes, stats, mbs};
    if (threads == 1) {
      slice_worker(&jobs[t]);
    } else if (pthread_create(&tids[t], NULL, slice_worker, &jobs[t]) == 0) {
      started++;
    } else {
      rc = -2;  /* the slices of this worker are left uncoded */
      break;
    }
  }
  for (int t = 0; t < started; t++) pthread_join(tids[t], NULL);
  for (size_t i = 0; i < ylen; i++) recon_y[i] = (uint8_t)pic[0][i];
  for (size_t i = 0; i < clen; i++) { recon_cb[i] = (uint8_t)pic[1][i]; recon_cr[i] = (uint8_t)pic[2][i]; }
  for (int r = 0; r < np; r++) {
    free(refs[r].hp);
    free((void *)refs[r].planes[1]);
    free((void *)refs[r].planes[2]);
  }
  for (int k = 0; k < 3; k++) { free(src[k]); free(pic[k]); }
  free(state);
  return rc;
}

int or_analyse_picture_mt(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                          int kind, const uint8_t *const *ref_y, const uint8_t *const *ref_cb,
                          const uint8_t *const *ref_cr, const ia_mbinfo *const *ref_mbs, int nref,
                          const ia_fsresult *fs, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y,
                          uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs, int threads) {
  if (kind == IA_PIC_B) return -1;
  return analyse_impl(cfg, y, cb, cr, qp, kind, ref_y, ref_cb, ref_cr, ref_mbs, nref, 0, 0, fs, codes, stats, recon_y,
                      recon_cb, recon_cr, mbs, threads);
}

int or_analyse_picture_b(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                         const uint8_t *const *ref_y, const uint8_t *const *ref_cb, const uint8_t *const *ref_cr,
                         const ia_mbinfo *const *ref_mbs, int tb, int td, ia_mbcode *codes, ia_mbstats *stats,
                         uint8_t *recon_y, uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs, int threads) {
  return analyse_impl(cfg, y, cb, cr, qp, IA_PIC_B, ref_y, ref_cb, ref_cr, ref_mbs, 2, tb, td, NULL, codes, stats,
                      recon_y, recon_cb, recon_cr, mbs, threads);
}

int or_analyse_picture(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                       int kind, const uint8_t *const *ref_y, const uint8_t *const *ref_cb,
                       const uint8_t *const *ref_cr, const ia_mbinfo *const *ref_mbs, int nref,
                       const ia_fsresult *fs, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y,
                       uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs) {
  return or_analyse_picture_mt(cfg, y, cb, cr, qp, kind, ref_y, ref_cb, ref_cr, ref_mbs, nref, fs, codes, stats,
                               recon_y, recon_cb, recon_cr, mbs, 1);
}

/* ---------------------------------------------------------------------------------------------
 * The loop filter as the encoder runs it on each reconstructed picture before it becomes a reference
 * (enc/mod.rs:653-654): frame macroblocks (no MBAFF), 4:2:0, 8-bit samples, every slice with
 * disable_deblocking_filter_idc 0, zero alpha/beta offsets and zero chroma QP offset; with per-slice
 * settings (the decoder's pictures) any idc, offsets and chroma QP offsets. I_PCM macroblocks carry
 * QP 0 in their records; lossless macroblocks are not handled. Follows h264/deblock.rs:232-421 and
 * :608-695.
 * ------------------------------------------------------------------------------------------- */

/* Whether the transform block holding 4x4 block (bx, by) has coefficients (h264/deblock.rs:336-344). */
static int db_coded(const ia_mbinfo *m, int bx, int by) {
  if (m->t8x8) {
    const int x0 = bx & 2, y0 = by & 2;
    return m->nnz[y0 * 4 + x0] | m->nnz[y0 * 4 + x0 + 1] | m->nnz[(y0 + 1) * 4 + x0] | m->nnz[(y0 + 1) * 4 + x0 + 1];
  }
  return m->nnz[by * 4 + bx] > 0;
}

/* Vectors a whole sample or more apart (h264/deblock.rs:347-349, frame macroblocks: mvy = 4). */
static int db_far(const int16_t *a, const int16_t *b) {
  return abs(a[0] - b[0]) >= 4 || abs(a[1] - b[1]) >= 4;
}

/* Different references, a different number of vectors, or vectors a sample apart
 * (h264/deblock.rs:352-383); list 1 is unused in P pictures (reference -1, vector 0). */
static int db_motion_differs(const ia_mbinfo *p, int pbx, int pby, const ia_mbinfo *q, int qbx, int qby) {
  const int p8 = (pby >> 1) * 2 + (pbx >> 1), q8 = (qby >> 1) * 2 + (qbx >> 1);
  const int32_t pr[2] = {p->ref_id[p8], p->ref_id1[p8]}, qr[2] = {q->ref_id[q8], q->ref_id1[q8]};
  const int16_t *pm[2] = {p->mv[pby * 4 + pbx], p->mv1[pby * 4 + pbx]};
  const int16_t *qm[2] = {q->mv[qby * 4 + qbx], q->mv1[qby * 4 + qbx]};
  const int np = (pr[0] >= 0) + (pr[1] >= 0), nq = (qr[0] >= 0) + (qr[1] >= 0);
  if (np != nq) return 1;
  if (np == 1) {
    const int lp = pr[0] >= 0 ? 0 : 1, lq = qr[0] >= 0 ? 0 : 1;
    return pr[lp] != qr[lq] || db_far(pm[lp], qm[lq]);
  }
  if (!((pr[0] == qr[0] && pr[1] == qr[1]) || (pr[0] == qr[1] && pr[1] == qr[0]))) return 1;
  if (pr[0] != pr[1]) {
    if (pr[0] == qr[0]) return db_far(pm[0], qm[0]) || db_far(pm[1], qm[1]);
    return db_far(pm[0], qm[1]) || db_far(pm[1], qm[0]);
  }
  return (db_far(pm[0], qm[0]) || db_far(pm[1], qm[1])) && (db_far(pm[0], qm[1]) || db_far(pm[1], qm[0]));
}

/* Every 4x4 block with the macroblock's first vector and every 8x8 block with its first reference,
 * in both lists (h264/deblock.rs:385-389). */
static int db_uniform_motion(const ia_mbinfo *m) {
  for (int i = 1; i < 4; i++)
    if (m->ref_idx[i] != m->ref_idx[0] || m->ref_id[i] != m->ref_id[0] || m->ref_idx1[i] != m->ref_idx1[0] ||
        m->ref_id1[i] != m->ref_id1[0])
      return 0;
  for (int i = 1; i < 16; i++)
    if (m->mv[i][0] != m->mv[0][0] || m->mv[i][1] != m->mv[0][1] || m->mv1[i][0] != m->mv1[0][0] ||
        m->mv1[i][1] != m->mv1[0][1])
      return 0;
  return 1;
}

/* Boundary strengths of the four 4-sample segments of edge e (0 = the macroblock edge) in direction
 * dir (0: vertical edges, 1: horizontal); `o` is the macroblock across a macroblock edge
 * (h264/deblock.rs:391-408). */
static void db_strength(const ia_mbinfo *m, const ia_mbinfo *o, int dir, int e, uint8_t *bs) {
  if (m->kind != IA_KIND_INTER || o->kind != IA_KIND_INTER) {
    for (int i = 0; i < 4; i++) bs[i] = e == 0 ? 4 : 3;
    return;
  }
  for (int i = 0; i < 4; i++) {
    int qbx, qby, pbx, pby;
    if (dir == 0) qbx = e, qby = i, pbx = e == 0 ? 3 : e - 1, pby = i;
    else qbx = i, qby = e, pbx = i, pby = e == 0 ? 3 : e - 1;
    bs[i] = db_coded(m, qbx, qby) || db_coded(o, pbx, pby) ? 2 : db_motion_differs(o, pbx, pby, m, qbx, qby) ? 1 : 0;
  }
}

/* One sample line across an edge; q is the first sample past it, xs the step across it
 * (h264/deblock.rs:279-334, 8-bit). */
static void db_filter_line(uint8_t *pl, ptrdiff_t q, ptrdiff_t xs, int b, int tc0, int alpha, int beta, int chroma) {
  const int p0 = pl[q - xs], q0 = pl[q], p1 = pl[q - 2 * xs], q1 = pl[q + xs];
  if (abs(p0 - q0) >= alpha || abs(p1 - p0) >= beta || abs(q1 - q0) >= beta) return;
  if (chroma) {
    if (b < 4) {
      const int tc = tc0 + 1;
      const int delta = clampi((((q0 - p0) * 4) + (p1 - q1) + 4) >> 3, -tc, tc);
      pl[q - xs] = (uint8_t)clampi(p0 + delta, 0, 255);
      pl[q] = (uint8_t)clampi(q0 - delta, 0, 255);
    } else {
      pl[q - xs] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
      pl[q] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
    }
    return;
  }
  const int p2 = pl[q - 3 * xs], q2 = pl[q + 2 * xs];
  const int ap = abs(p2 - p0), aq = abs(q2 - q0);
  if (b < 4) {
    const int tc = tc0 + (ap < beta) + (aq < beta);
    const int delta = clampi((((q0 - p0) * 4) + (p1 - q1) + 4) >> 3, -tc, tc);
    pl[q - xs] = (uint8_t)clampi(p0 + delta, 0, 255);
    pl[q] = (uint8_t)clampi(q0 - delta, 0, 255);
    if (ap < beta) pl[q - 2 * xs] = (uint8_t)(p1 + clampi((p2 + ((p0 + q0 + 1) >> 1) - (2 * p1)) >> 1, -tc0, tc0));
    if (aq < beta) pl[q + xs] = (uint8_t)(q1 + clampi((q2 + ((p0 + q0 + 1) >> 1) - (2 * q1)) >> 1, -tc0, tc0));
    return;
  }
  const int small = abs(p0 - q0) < (alpha >> 2) + 2;
  if (ap < beta && small) {
    const int p3 = pl[q - 4 * xs];
    pl[q - xs] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
    pl[q - 2 * xs] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
    pl[q - 3 * xs] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
  } else {
    pl[q - xs] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
  }
  if (aq < beta && small) {
    const int q3 = pl[q + 3 * xs];
    pl[q] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
    pl[q + xs] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
    pl[q + 2 * xs] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
  } else {
    pl[q] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
  }
}

/* An edge of segments bs[0..3] of `per_bs` lines each, starting at pos, lines ys apart
 * (h264/deblock.rs:232-276 without I_PCM restoring). */
static void db_filter_edge(uint8_t *pl, ptrdiff_t pos, ptrdiff_t xs, ptrdiff_t ys, int per_bs, const uint8_t *bs,
                           int qpav, int chroma, int alpha_offset, int beta_offset) {
  const int ia = clampi(qpav + alpha_offset, 0, 51), ib = clampi(qpav + beta_offset, 0, 51);
  const int alpha = OR_ALPHA[ia], beta = OR_BETA[ib];
  if (alpha == 0 || beta == 0) return;
  for (int s = 0; s < 4; s++) {
    if (bs[s] == 0) continue;
    const int tc0 = bs[s] < 4 ? (OR_TC0[ia][bs[s]] > 0 ? OR_TC0[ia][bs[s]] : 0) : 0;
    for (int k = s * per_bs; k < (s + 1) * per_bs; k++) db_filter_line(pl, pos + k * ys, xs, bs[s], tc0, alpha, beta, chroma);
  }
}

/* Macroblock (mbx, mby): strengths once, then luma and chroma edges (h264/deblock.rs:608-695). */
static void db_mb(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mbx, int mby,
                  const or_slice_filter *sp) {
  const ptrdiff_t ls = (ptrdiff_t)mb_w * 16, cs = (ptrdiff_t)mb_w * 8;
  const int idx = mby * mb_w + mbx;
  const ia_mbinfo *m = &mbs[idx];
  if (m->slice == 0) return;
  /* the filter settings of the slice holding this macroblock (h264/deblock.rs:608-620): idc 1 leaves
     its edges alone, idc 2 keeps to edges inside the slice, idc 0 filters across slice edges too */
  const or_slice_filter zero = {0, 0, 0, {0, 0}};
  const or_slice_filter *f = sp ? &sp[m->slice - 1] : &zero;
  if (f->disable_idc == 1) return;
  const int left_ok = mbx > 0 && mbs[idx - 1].slice != 0 && (f->disable_idc != 2 || mbs[idx - 1].slice == m->slice);
  const int top_ok = mby > 0 && mbs[idx - mb_w].slice != 0 && (f->disable_idc != 2 || mbs[idx - mb_w].slice == m->slice);
  const ia_mbinfo *edge_mb[2] = {mbx > 0 ? &mbs[idx - 1] : m, mby > 0 ? &mbs[idx - mb_w] : m};
  uint8_t bs[2][4][4];
  memset(bs, 0, sizeof bs);
  int still = m->kind == IA_KIND_INTER && db_uniform_motion(m);
  for (int i = 0; i < 16 && still; i++) still = m->nnz[i] == 0;
  for (int dir = 0; dir < 2; dir++)
    for (int e = 0; e < 4; e++) {
      if (e > 0 && still) continue;
      db_strength(m, e == 0 ? edge_mb[dir] : m, dir, e, bs[dir][e]);
    }
  const int x0 = mbx * 16, y0 = mby * 16;
  for (int dir = 0; dir < 2; dir++)
    for (int e = 0; e < 4; e++) {
      if (e == 0 && !(dir == 0 ? left_ok : top_ok)) continue;
      if ((m->t8x8 && (e & 1)) || !(bs[dir][e][0] | bs[dir][e][1] | bs[dir][e][2] | bs[dir][e][3])) continue;
      const ia_mbinfo *o = e == 0 ? edge_mb[dir] : m;
      const int qpav = (o->qp + m->qp + 1) >> 1;
      if (dir == 0) db_filter_edge(y, y0 * ls + x0 + e * 4, 1, ls, 4, bs[dir][e], qpav, 0, f->alpha_offset, f->beta_offset);
      else db_filter_edge(y, (y0 + e * 4) * ls + x0, ls, 1, 4, bs[dir][e], qpav, 0, f->alpha_offset, f->beta_offset);
    }
  /* chroma 4:2:0: vertical edges 0 and 4, horizontal edges 0 and 4 (luma edges 0 and 2) */
  uint8_t *cpl[2] = {cb, cr};
  for (int comp = 0; comp < 2; comp++)
    for (int dir = 0; dir < 2; dir++)
      for (int e = 0; e < 4; e += 2) {
        if ((e == 0 && !(dir == 0 ? left_ok : top_ok)) || !(bs[dir][e][0] | bs[dir][e][1] | bs[dir][e][2] | bs[dir][e][3]))
          continue;
        const ia_mbinfo *o = e == 0 ? edge_mb[dir] : m;
        const int off = f->chroma_qp_offset[comp];
        const int qpav = (chroma_qp(o->qp, off, 8) + chroma_qp(m->qp, off, 8) + 1) >> 1;
        const int cx0 = mbx * 8, cy0 = mby * 8;
        if (dir == 0)
          db_filter_edge(cpl[comp], cy0 * cs + cx0 + e * 2, 1, cs, 2, bs[dir][e], qpav, 1, f->alpha_offset, f->beta_offset);
        else
          db_filter_edge(cpl[comp], (cy0 + e * 2) * cs + cx0, cs, 1, 2, bs[dir][e], qpav, 1, f->alpha_offset, f->beta_offset);
      }
}

void or_deblock_picture(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h) {
  or_deblock_picture_params(y, cb, cr, mbs, mb_w, mb_h, NULL);
}

void or_deblock_picture_params(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h,
                               const or_slice_filter *sp) {
  for (int mby = 0; mby < mb_h; mby++)
    for (int mbx = 0; mbx < mb_w; mbx++) db_mb(y, cb, cr, mbs, mb_w, mbx, mby, sp);
}

/* Sizes of the ABI structs as this compiler lays them out, for the binding tests. */
void or_abi_sizes(int32_t *out) {
  out[0] = (int32_t)sizeof(ia_config);
  out[1] = (int32_t)sizeof(ia_mbinfo);
  out[2] = (int32_t)sizeof(ia_mbcode);
  out[3] = (int32_t)sizeof(ia_mbstats);
  out[4] = (int32_t)sizeof(ia_fsresult);
  out[5] = (int32_t)sizeof(ia_frame);
}
