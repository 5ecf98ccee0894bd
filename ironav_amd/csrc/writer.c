/* writer.c — the bitstream back end on the host: the analysed macroblock codes of a picture written as
 * an H.264 access unit (Annex B), as nervosys/IronAV's encoder writes them around its slice coders.
 *
 * Restated from the reference encoder:
 *  - NAL units with emulation prevention and Exp-Golomb codes: `BitWriter`, `put_nal`
 *    (crates/ironav-codec/src/h264/enc/bitw.rs:4-105);
 *  - the CABAC arithmetic coder: `CabacWriter` (enc/cabac.rs:42-146), contexts initialised per slice
 *    (enc/cabac.rs:54-64);
 *  - the macroblock bins: `write_mb_cabac` and `mb_bins` with `write_block` and the mvd bins
 *    (enc/frame.rs:2020-2279, enc/cabac.rs:194-246, enc/frame.rs:2393-2413), and the neighbour state
 *    `finish_mb` keeps for the next macroblocks (enc/frame.rs:2281-2339);
 *  - the parameter sets and slice headers of `code_picture` for IPPP coding: `sps`, `pps` and the slice
 *    header closure (enc/mod.rs:296-400, :553-603), the level from Table A-1 as `H264Encoder::new`
 *    picks it (enc/mod.rs:43-64, :155-158), the slice bands (enc/mod.rs:603) and NAL order
 *    (enc/mod.rs:636-652).
 * CABAC only: the CAVLC writer (enc/cavlc.rs, `write_mb`) is not restated.
 */
#include <stdlib.h>
#include <string.h>

#include "../../include/ironav_h264w.h"
#include "h264_tables.h"

enum { K_NONE = 0, K_I4 = 1, K_I8 = 2, K_I16 = 3, K_INTER = 5 };

static inline int mini(int a, int b) { return a < b ? a : b; }
static inline int absi(int a) { return a < 0 ? -a : a; }

/* ---- bits (enc/bitw.rs) ------------------------------------------------------------------------- */

typedef struct {
  uint8_t *buf;
  size_t len, cap;
  uint64_t cur;
  int n;
  int overflow;
} bw_t;

static void bw_byte(bw_t *w, uint8_t b) {
  if (w->len >= w->cap) {
    size_t nc = w->cap ? 2 * w->cap : 4096;
    uint8_t *nb = (uint8_t *)realloc(w->buf, nc);
    if (!nb) {
      w->overflow = 1;
      return;
    }
    w->buf = nb;
    w->cap = nc;
  }
  w->buf[w->len++] = b;
}

static void bw_put(bw_t *w, int n, uint32_t v) {
  w->cur = (w->cur << n) | v;
  w->n += n;
  while (w->n >= 8) {
    w->n -= 8;
    bw_byte(w, (uint8_t)(w->cur >> w->n));
  }
  w->cur &= (1ull << w->n) - 1;
}

static void bw_flag(bw_t *w, int b) { bw_put(w, 1, b ? 1u : 0u); }

static void bw_ue(bw_t *w, uint32_t v) {
  const uint64_t x = (uint64_t)v + 1;
  const int len = 64 - __builtin_clzll(x);
  if (len > 16) {
    bw_put(w, len - 1, 0);
    bw_put(w, len, (uint32_t)x);
  } else {
    bw_put(w, 2 * len - 1, (uint32_t)x);
  }
}

static void bw_se(bw_t *w, int32_t v) { bw_ue(w, v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * (int64_t)v)); }

static void bw_align_ones(bw_t *w) {
  if (w->n > 0) {
    const int pad = 8 - w->n;
    bw_put(w, pad, (1u << pad) - 1);
  }
}

static void bw_pad_zeros(bw_t *w) {
  if (w->n > 0) bw_put(w, 8 - w->n, 0);
}

/* rbsp_trailing_bits */
static void bw_finish(bw_t *w) {
  bw_put(w, 1, 1);
  bw_pad_zeros(w);
}

int iw_put_nal(int ref_idc, int type, const uint8_t *rbsp, size_t n, uint8_t *out, size_t cap, size_t *len) {
  if (!rbsp || !len || ref_idc < 0 || ref_idc > 3 || type < 0 || type > 31) return IW_EINVAL;
  size_t o = 0;
  const uint8_t head[5] = {0, 0, 0, 1, (uint8_t)(ref_idc << 5 | type)};
#define PUT(b)                   \
  do {                           \
    if (out && o < cap) out[o] = (b); \
    o++;                         \
  } while (0)
  for (int i = 0; i < 5; i++) PUT(head[i]);
  int zeros = 0;
  for (size_t i = 0; i < n; i++) {
    const uint8_t b = rbsp[i];
    if (zeros >= 2 && b <= 3) {
      PUT(3);
      zeros = 0;
    }
    PUT(b);
    zeros = b == 0 ? zeros + 1 : 0;
  }
#undef PUT
  *len = o;
  return out && o > cap ? IW_ESPACE : IW_OK;
}

/* ---- the CABAC coder (enc/cabac.rs:42-146) ------------------------------------------------------- */

typedef struct {
  bw_t w;
  uint32_t low, range, outstanding;
  int first;
  uint8_t st[1024]; /* pStateIdx << 1 | valMPS */
} cw_t;

static void cw_init_contexts(cw_t *c, int intra, int qp) {
  qp = qp < 0 ? 0 : (qp > 51 ? 51 : qp);
  for (int i = 0; i < 1024; i++) {
    const int m = intra ? IA_CABAC_INIT_I[i][0] : IA_CABAC_INIT_P[i][0];
    const int n = intra ? IA_CABAC_INIT_I[i][1] : IA_CABAC_INIT_P[i][1];
    int pre = ((m * qp) >> 4) + n;
    pre = pre < 1 ? 1 : (pre > 126 ? 126 : pre);
    c->st[i] = pre <= 63 ? (uint8_t)((63 - pre) << 1) : (uint8_t)(((pre - 64) << 1) | 1);
  }
  c->low = 0;
  c->range = 510;
  c->outstanding = 0;
  c->first = 1;
}

static void cw_put_bit(cw_t *c, uint32_t b) {
  if (c->first) c->first = 0;
  else bw_put(&c->w, 1, b);
  while (c->outstanding > 0) {
    bw_put(&c->w, 1, 1 - b);
    c->outstanding--;
  }
}

static void cw_renorm(cw_t *c) {
  while (c->range < 256) {
    if (c->low < 256) {
      cw_put_bit(c, 0);
    } else if (c->low >= 512) {
      c->low -= 512;
      cw_put_bit(c, 1);
    } else {
      c->low -= 256;
      c->outstanding++;
    }
    c->range <<= 1;
    c->low <<= 1;
  }
}

static void cw_decision(cw_t *c, int ctx, int bin) {
  const int s = c->st[ctx], p = s >> 1, mps = s & 1;
  const uint32_t lps = IA_LPS_RANGE[p][(c->range >> 6) & 3];
  c->range -= lps;
  if (bin != mps) {
    c->low += c->range;
    c->range = lps;
    c->st[ctx] = (uint8_t)((IA_TRANS_LPS[p] << 1) | (p == 0 ? !mps : mps));
  } else {
    c->st[ctx] = (uint8_t)(((p < 62 ? p + 1 : p) << 1) | mps);
  }
  cw_renorm(c);
}

static void cw_bypass(cw_t *c, int bin) {
  c->low <<= 1;
  if (bin) c->low += c->range;
  if (c->low >= 1024) {
    cw_put_bit(c, 1);
    c->low -= 1024;
  } else if (c->low < 512) {
    cw_put_bit(c, 0);
  } else {
    c->low -= 512;
    c->outstanding++;
  }
}

static void cw_terminate(cw_t *c, int bin) {
  c->range -= 2;
  if (bin) {
    c->low += c->range;
    c->range = 2;
    cw_renorm(c);
    cw_put_bit(c, (c->low >> 9) & 1);
    /* the last of these two bits is the rbsp_stop_one_bit */
    bw_put(&c->w, 2, ((c->low >> 7) & 3) | 1);
  } else {
    cw_renorm(c);
  }
}

static void cw_exp_golomb(cw_t *c, uint32_t v, int k) {
  while (v >= (1u << k)) {
    cw_bypass(c, 1);
    v -= 1u << k;
    k++;
  }
  cw_bypass(c, 0);
  while (k > 0) {
    k--;
    cw_bypass(c, (v >> k) & 1);
  }
}

int iw_cabac_encode_bins(const uint8_t *kinds, const uint16_t *ctx, const uint8_t *vals, int n, int intra, int qp,
                         uint8_t *out, size_t cap, size_t *len) {
  if (!kinds || !ctx || !vals || n < 0 || !len) return IW_EINVAL;
  cw_t c;
  memset(&c, 0, sizeof c);
  cw_init_contexts(&c, intra, qp);
  for (int i = 0; i < n; i++) {
    if (kinds[i] == 0) {
      if (ctx[i] >= 1024) {
        free(c.w.buf);
        return IW_EINVAL;
      }
      cw_decision(&c, ctx[i], vals[i] != 0);
    } else if (kinds[i] == 1) {
      cw_bypass(&c, vals[i] != 0);
    } else {
      cw_terminate(&c, 0);
    }
  }
  cw_terminate(&c, 1);
  bw_pad_zeros(&c.w);
  int rc = c.w.overflow ? IW_ENOMEM : IW_OK;
  *len = c.w.len;
  if (!rc && out) {
    if (c.w.len > cap) rc = IW_ESPACE;
    else memcpy(out, c.w.buf, c.w.len);
  }
  free(c.w.buf);
  return rc;
}

/* ---- macroblock bins (enc/frame.rs:2028-2279) ---------------------------------------------------- */

/* what the bins of later macroblocks read about a written one (enc/frame.rs:239-272) */
typedef struct {
  uint8_t i4, inter, t8x8, skip, kind, cbp, chroma_mode, cbf_dc;
  uint8_t modes[16], nnz[16], cnnz[2][4];
  int8_t refs[4];
  uint8_t bdirect;                /* B_Skip or B_Direct_16x16 */
  uint8_t mvd[16][2], mvd1[16][2]; /* lists 0 and 1 */
} mbstate_t;

typedef struct {
  int x, y, w, h;
} rect_t;

static const uint8_t PARTS[4][4][4] = {{{0, 0, 4, 4}}, {{0, 0, 4, 2}, {0, 2, 4, 2}}, {{0, 0, 2, 4}, {2, 0, 2, 4}},
                                       {{0, 0, 2, 2}, {2, 0, 2, 2}, {0, 2, 2, 2}, {2, 2, 2, 2}}};
static const int NPARTS[4] = {1, 2, 2, 4};
static const uint8_t SUBS[4][4][4] = {{{0, 0, 2, 2}}, {{0, 0, 2, 1}, {0, 1, 2, 1}}, {{0, 0, 1, 2}, {1, 0, 1, 2}},
                                      {{0, 0, 1, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}}};
static const int NSUBS[4] = {1, 2, 2, 4};

/* the partitions of a macroblock in coding order, in 4x4 blocks (enc/frame.rs:182-236 `rects`) */
static int rects(int part, const uint8_t *sub, rect_t *out) {
  int n = 0;
  if (part != 3) {
    for (int k = 0; k < NPARTS[part]; k++)
      out[n++] = (rect_t){PARTS[part][k][0], PARTS[part][k][1], PARTS[part][k][2], PARTS[part][k][3]};
    return n;
  }
  for (int k = 0; k < 4; k++) {
    const int s = sub[k] & 3;
    for (int q = 0; q < NSUBS[s]; q++)
      out[n++] = (rect_t){(k & 1) * 2 + SUBS[s][q][0], (k >> 1) * 2 + SUBS[s][q][1], SUBS[s][q][2], SUBS[s][q][3]};
  }
  return n;
}

static void block_pos(int n, int *bx, int *by) {
  *bx = ((n >> 2) & 1) * 2 + (n & 1);
  *by = ((n >> 3) & 1) * 2 + ((n >> 1) & 1);
}

static void mvd_grid(const rect_t *rs, int n, const int16_t (*mvds)[2], uint8_t (*out)[2]) {
  memset(out, 0, 32);
  for (int k = 0; k < n; k++)
    for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
      for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++)
        for (int q = 0; q < 2; q++) out[by * 4 + bx][q] = (uint8_t)mini(absi(mvds[k][q]), 64);
}

static int count_nz(const int16_t *v, int n) {
  int k = 0;
  for (int i = 0; i < n; i++) k += v[i] != 0;
  return k;
}

/* enc/frame.rs:2281-2339, the neighbour state; b: a macroblock of a B slice */
static void finish_state(const ia_mbcode *code, mbstate_t *s, int b) {
  const int inter = code->kind == K_INTER;
  static const uint8_t Z4[4] = {0, 0, 0, 0};
  memset(s, 0, sizeof *s);
  rect_t rs[16];
  const int n = rects(code->part & 3, b ? Z4 : code->sub, rs);
  if (inter && !code->skip) {
    if (b) {
      mvd_grid(rs, n, (const int16_t(*)[2])code->bmvd[0], s->mvd);
      mvd_grid(rs, n, (const int16_t(*)[2])code->bmvd[1], s->mvd1);
    } else {
      mvd_grid(rs, n, (const int16_t(*)[2])code->mvds, s->mvd);
    }
  }
  s->bdirect = (uint8_t)(b && inter && (code->part & 3) == 0 && code->bdir[0] == 0);
  const int cbpc = code->cbp >> 4;
  s->cbf_dc = (uint8_t)((code->kind == K_I16 && count_nz(code->dc, 16) > 0) |
                        ((cbpc > 0 && count_nz(code->cdc[0], 4) > 0) << 1) | ((cbpc > 0 && count_nz(code->cdc[1], 4) > 0) << 2));
  int8_t r16[16];
  const int part = code->part & 3;
  for (int k = 0; k < NPARTS[part]; k++) {
    const uint8_t *pr = PARTS[part][k];
    for (int by = pr[1]; by < pr[1] + pr[3]; by++)
      for (int bx = pr[0]; bx < pr[0] + pr[2]; bx++) r16[by * 4 + bx] = (int8_t)code->refs[k];
  }
  s->i4 = code->kind == K_I4 || code->kind == K_I8;
  s->t8x8 = code->t8x8;
  s->inter = (uint8_t)inter;
  memcpy(s->modes, code->modes, 16);
  memcpy(s->nnz, code->nnz, 16);
  memcpy(s->cnnz, code->cnnz, 8);
  for (int b8 = 0; b8 < 4; b8++) s->refs[b8] = inter ? r16[(b8 / 2) * 8 + (b8 % 2) * 2] : -1;
  s->skip = code->skip;
  s->kind = code->kind;
  s->cbp = code->cbp;
  s->chroma_mode = inter ? 0 : code->chroma_mode;
}

/* Intra4x4/8x8 mode predictor from the modes coded so far in the macroblock (enc/frame.rs:619-627) */
static int pred_mode(const uint8_t *cur, int bx, int by, const mbstate_t *left, const mbstate_t *top) {
  int a = -1, b = -1;
  if (bx > 0) a = cur[by * 4 + bx - 1];
  else if (left) a = left->i4 ? left->modes[by * 4 + 3] : 2;
  if (by > 0) b = cur[(by - 1) * 4 + bx];
  else if (top) b = top->i4 ? top->modes[12 + bx] : 2;
  if (a >= 0 && b >= 0) return a < b ? a : b;
  return 2;
}

/* residual_block_cabac (enc/cabac.rs:194-246) */
static void write_block(cw_t *c, int cat, int cbf_ctx, const int16_t *levels, int max) {
  int total = 0, last = -1;
  for (int i = 0; i < max; i++)
    if (levels[i]) {
      total++;
      last = i;
    }
  if (cbf_ctx >= 0) cw_decision(c, cbf_ctx, total > 0);
  if (total == 0) return;
  for (int i = 0; i < max - 1; i++) {
    int sig, lst;
    if (cat == 3) sig = lst = mini(i, 2);
    else if (cat == 5) {
      sig = IA_SIG_OFF_8X8[0][i];
      lst = IA_LAST_OFF_8X8[i];
    } else sig = lst = i;
    const int nz = levels[i] != 0;
    cw_decision(c, IA_SIG_OFF[0][cat] + sig, nz);
    if (nz) {
      cw_decision(c, IA_LAST_OFF[0][cat] + lst, i == last);
      if (i == last) break;
    }
  }
  const int base = IA_ABS_OFF[cat];
  int eq1 = 0, gt1 = 0;
  for (int i = last; i >= 0; i--) {
    const int l = levels[i];
    if (!l) continue;
    const int a = absi(l);
    const int inc0 = gt1 != 0 ? 0 : mini(1 + eq1, 4);
    if (a == 1) {
      cw_decision(c, base + inc0, 0);
      eq1++;
    } else {
      cw_decision(c, base + inc0, 1);
      const int incn = base + 5 + mini(gt1, 4 - (cat == 3));
      for (int k = 0; k < mini(a - 2, 13); k++) cw_decision(c, incn, 1);
      if (a < 15) cw_decision(c, incn, 0);
      else cw_exp_golomb(c, (uint32_t)(a - 15), 0);
      gt1++;
    }
    cw_bypass(c, l < 0);
  }
}

/* enc/frame.rs:2393-2413 */
static void mvd_bins(cw_t *c, int comp, int d, int sum) {
  const int base = comp == 0 ? 40 : 47;
  const int inc = sum < 3 ? 0 : (sum > 32 ? 2 : 1);
  const int a = absi(d);
  cw_decision(c, base + inc, a > 0);
  if (a > 0) {
    int ctx = base + 3;
    for (int k = 1; k < mini(a, 9); k++) {
      cw_decision(c, ctx, 1);
      if (ctx < base + 6) ctx++;
    }
    if (a < 9) cw_decision(c, ctx, 0);
    else cw_exp_golomb(c, (uint32_t)(a - 9), 3);
    cw_bypass(c, d < 0);
  }
}

/* The mb_type of a B macroblock (Table 7-14; enc/frame.rs:2343-2351). */
static int b_mb_type(const ia_mbcode *code) {
  static const uint8_t PAIRS[9][2] = {{1, 1}, {2, 2}, {1, 2}, {2, 1}, {1, 3}, {2, 3}, {3, 1}, {3, 2}, {3, 3}};
  const int part = code->part & 3;
  if (part == 0) return code->bdir[0];
  if (part == 3) return 22;
  int i = 0;
  while (i < 8 && !(PAIRS[i][0] == code->bdir[0] && PAIRS[i][1] == code->bdir[1])) i++;
  return 4 + 2 * i + (part == 2);
}

static void mb_bins(cw_t *c, const ia_mbcode *code, const mbstate_t *left, const mbstate_t *top, int p_slice,
                    int b_slice, int transform8, int nref) {
  const int kind = code->kind;
  if (p_slice) {
    cw_decision(c, (b_slice ? 24 : 11) + (left && !left->skip) + (top && !top->skip), code->skip);
    if (code->skip) return;
  }
  const int cbpl = code->cbp & 15, cbpc = code->cbp >> 4, prefix_ok = p_slice, base = b_slice ? 32 : p_slice ? 17 : 3;
  const int part = code->part & 3;
  const int b_ctx = 27 + (left && !left->bdirect) + (top && !top->bdirect);
  if (kind == K_INTER && b_slice) {
    /* B_Direct_16x16 0, B_L0_16x16 100, B_L1_16x16 101, the others 11 and four or five more bins
       (Table 9-37) */
    const int t = b_mb_type(code);
    cw_decision(c, b_ctx, t != 0);
    if (t == 1 || t == 2) {
      cw_decision(c, 30, 0);
      cw_decision(c, 32, t == 2);
    } else if (t != 0) {
      cw_decision(c, 30, 1);
      const int bits = t >= 3 && t <= 10 ? t - 3 : t == 11 ? 14 : t == 22 ? 15 : (t + 4) >> 1;
      cw_decision(c, 31, (bits >> 3) & 1);
      for (int sh = 2; sh >= 0; sh--) cw_decision(c, 32, (bits >> sh) & 1);
      if (t >= 12 && t <= 21) cw_decision(c, 32, (t + 4) & 1);
    }
    if (t == 22)
      /* sub_mb_type: B_Direct_8x8 0, B_L0_8x8 100, B_L1_8x8 101, B_Bi_8x8 11000 */
      for (int k = 0; k < 4; k++) {
        const int d = code->bdir[k];
        cw_decision(c, 36, d != 0);
        if (d == 1 || d == 2) {
          cw_decision(c, 37, 0);
          cw_decision(c, 39, d == 2);
        } else if (d == 3) {
          cw_decision(c, 37, 1);
          cw_decision(c, 38, 0);
          cw_decision(c, 39, 0);
          cw_decision(c, 39, 0);
        }
      }
  } else if (kind == K_INTER) {
    const int b1 = part == 1 || part == 2, b2 = part == 1 || part == 3;
    cw_decision(c, 14, 0);
    cw_decision(c, 15, b1);
    cw_decision(c, b1 ? 17 : 16, b2);
  } else {
    int first;
    if (b_slice) {
      /* 111101: an intra type follows, with the suffix contexts */
      cw_decision(c, b_ctx, 1);
      cw_decision(c, 30, 1);
      cw_decision(c, 31, 1);
      cw_decision(c, 32, 1);
      cw_decision(c, 32, 0);
      cw_decision(c, 32, 1);
      first = 32;
    } else if (prefix_ok) {
      cw_decision(c, 14, 1);
      first = 17;
    } else {
      first = 3 + (left && !(left->kind == K_I4 || left->kind == K_I8)) + (top && !(top->kind == K_I4 || top->kind == K_I8));
    }
    if (kind == K_I4 || kind == K_I8) {
      cw_decision(c, first, 0);
      if (transform8) cw_decision(c, 399 + (left && left->t8x8) + (top && top->t8x8), kind == K_I8);
    } else {
      cw_decision(c, first, 1);
      cw_terminate(c, 0);
      const int b1 = prefix_ok ? base + 1 : base + 3, b2 = prefix_ok ? base + 2 : base + 4, b3 = prefix_ok ? base + 3 : base + 6;
      cw_decision(c, b1, cbpl != 0);
      cw_decision(c, b2, cbpc != 0);
      if (cbpc != 0) cw_decision(c, prefix_ok ? base + 2 : base + 5, cbpc == 2);
      cw_decision(c, b3, (code->i16_mode >> 1) == 1);
      cw_decision(c, prefix_ok ? base + 3 : base + 7, (code->i16_mode & 1) == 1);
    }
  }
  if (kind == K_I4 || kind == K_I8) {
    const int t8 = kind == K_I8;
    uint8_t so_far[16];
    memset(so_far, 2, 16);
    for (int n = 0; n < (t8 ? 4 : 16); n++) {
      int bx, by;
      if (t8) {
        bx = (n & 1) * 2;
        by = (n >> 1) * 2;
      } else {
        block_pos(n, &bx, &by);
      }
      const int mode = code->modes[by * 4 + bx];
      const int pred = pred_mode(so_far, bx, by, left, top);
      so_far[by * 4 + bx] = (uint8_t)mode;
      if (t8) {
        so_far[by * 4 + bx + 1] = (uint8_t)mode;
        so_far[(by + 1) * 4 + bx] = (uint8_t)mode;
        so_far[(by + 1) * 4 + bx + 1] = (uint8_t)mode;
      }
      cw_decision(c, 68, mode == pred);
      if (mode != pred) {
        const int rem = mode < pred ? mode : mode - 1;
        for (int k = 0; k < 3; k++) cw_decision(c, 69, (rem >> k) & 1);
      }
    }
  }
  if (kind != K_INTER) {
    const int m = code->chroma_mode;
    const int ca = left && !left->inter && left->chroma_mode != 0, cb = top && !top->inter && top->chroma_mode != 0;
    cw_decision(c, 64 + ca + cb, m > 0);
    if (m > 0) {
      cw_decision(c, 67, m > 1);
      if (m > 1) cw_decision(c, 67, m > 2);
    }
  } else {
    if (!b_slice && part == 3)
      for (int k = 0; k < 4; k++) {
        const int s = code->sub[k];
        if (s == 0) cw_decision(c, 21, 1);
        else if (s == 1) {
          cw_decision(c, 21, 0);
          cw_decision(c, 22, 0);
        } else if (s == 2) {
          cw_decision(c, 21, 0);
          cw_decision(c, 22, 1);
          cw_decision(c, 23, 1);
        } else {
          cw_decision(c, 21, 0);
          cw_decision(c, 22, 1);
          cw_decision(c, 23, 0);
        }
      }
    if (!b_slice && nref > 1) {
      uint8_t grid[4] = {0, 0, 0, 0};
      rect_t parts[4];
      int np;
      if (part == 3) {
        np = 4;
        for (int k = 0; k < 4; k++) parts[k] = (rect_t){(k & 1) * 2, (k >> 1) * 2, 2, 2};
      } else {
        np = NPARTS[part];
        for (int k = 0; k < np; k++) parts[k] = (rect_t){PARTS[part][k][0], PARTS[part][k][1], PARTS[part][k][2], PARTS[part][k][3]};
      }
      for (int k = 0; k < np; k++) {
        const int bx = parts[k].x, by = parts[k].y, w = parts[k].w, h = parts[k].h;
        const int a = bx > 0 ? grid[(by / 2) * 2 + (bx - 1) / 2] > 0 : (left && left->refs[(by / 2) * 2 + 1] > 0);
        const int bb = by > 0 ? grid[((by - 1) / 2) * 2 + bx / 2] > 0 : (top && top->refs[2 + bx / 2] > 0);
        const int r = code->refs[k], cbase = 54 + a + 2 * bb;
        for (int i = 0; i <= r; i++) cw_decision(c, i == 0 ? cbase : (i == 1 ? 58 : 59), i < r);
        for (int y8 = by / 2; y8 <= (by + h - 1) / 2; y8++)
          for (int x8 = bx / 2; x8 <= (bx + w - 1) / 2; x8++) grid[y8 * 2 + x8] = code->refs[k];
      }
    }
    /* list 0's differences for the partitions using it, then list 1's */
    static const uint8_t Z4[4] = {0, 0, 0, 0};
    rect_t rs[16];
    const int n = rects(part, b_slice ? Z4 : code->sub, rs);
    for (int l = 0; l < (b_slice ? 2 : 1); l++) {
      const int16_t(*mvds)[2] = b_slice ? (const int16_t(*)[2])code->bmvd[l] : (const int16_t(*)[2])code->mvds;
      uint8_t grid[16][2];
      mvd_grid(rs, n, mvds, grid);
      for (int k = 0; k < n; k++) {
        if (b_slice && !((code->bdir[k] >> l) & 1)) continue;
        const int bx = rs[k].x, by = rs[k].y;
        for (int comp = 0; comp < 2; comp++) {
          const int a = bx > 0 ? grid[by * 4 + bx - 1][comp] : (left ? (l ? left->mvd1 : left->mvd)[by * 4 + 3][comp] : 0);
          const int bb = by > 0 ? grid[(by - 1) * 4 + bx][comp] : (top ? (l ? top->mvd1 : top->mvd)[12 + bx][comp] : 0);
          mvd_bins(c, comp, mvds[k][comp], a + bb);
        }
      }
    }
  }
  if (kind != K_I16) {
    int luma = 0;
    for (int b8 = 0; b8 < 4; b8++) {
      const int cond_a = (b8 & 1) ? ((luma >> (b8 - 1)) & 1) == 0 : (left && ((left->cbp >> (b8 + 1)) & 1) == 0);
      const int cond_b = b8 >= 2 ? ((luma >> (b8 - 2)) & 1) == 0 : (top && ((top->cbp >> (b8 + 2)) & 1) == 0);
      const int bit = (cbpl >> b8) & 1;
      cw_decision(c, 73 + cond_a + 2 * cond_b, bit);
      luma |= bit << b8;
    }
    int ca = left && (left->cbp >> 4) != 0, cb = top && (top->cbp >> 4) != 0;
    cw_decision(c, 77 + ca + 2 * cb, cbpc != 0);
    if (cbpc != 0) {
      ca = left && (left->cbp >> 4) == 2;
      cb = top && (top->cbp >> 4) == 2;
      cw_decision(c, 81 + ca + 2 * cb, cbpc == 2);
    }
  }
  if (kind == K_INTER && transform8 && cbpl != 0 && (part != 3 || (code->sub[0] | code->sub[1] | code->sub[2] | code->sub[3]) == 0))
    cw_decision(c, 399 + (left && left->t8x8) + (top && top->t8x8), code->t8x8);
  if (code->cbp != 0 || kind == K_I16) {
    cw_decision(c, 60, 0); /* mb_qp_delta 0: the slice QP throughout */
    const int un = kind != K_INTER;
    if (kind == K_I16) {
      const int da = left ? (left->cbf_dc & 1) != 0 : un, db = top ? (top->cbf_dc & 1) != 0 : un;
      write_block(c, 0, 85 + da + 2 * db, code->dc, 16);
    }
    const int cat = kind == K_I16 ? 1 : 2, cbase = kind == K_I16 ? 85 + 4 : 85 + 8;
    for (int i8 = 0; i8 < 4; i8++)
      if (code->t8x8 && (cbpl & (1 << i8))) write_block(c, 5, -1, code->luma8[i8], 64);
    for (int n = 0; n < 16; n++) {
      int bx, by;
      block_pos(n, &bx, &by);
      if (code->t8x8 || !(cbpl & (1 << (n / 4)))) continue;
      const int a = bx > 0 ? code->nnz[by * 4 + bx - 1] > 0 : (left ? left->nnz[by * 4 + 3] > 0 : un);
      const int bb = by > 0 ? code->nnz[(by - 1) * 4 + bx] > 0 : (top ? top->nnz[12 + bx] > 0 : un);
      const int16_t *blk = code->luma[by * 4 + bx];
      if (cat == 1) write_block(c, cat, cbase + a + 2 * bb, blk + 1, 15);
      else write_block(c, cat, cbase + a + 2 * bb, blk, 16);
    }
    if (cbpc > 0)
      for (int comp = 0; comp < 2; comp++) {
        const int da = left ? (left->cbf_dc & (2 << comp)) != 0 : un, db = top ? (top->cbf_dc & (2 << comp)) != 0 : un;
        write_block(c, 3, 85 + 12 + da + 2 * db, code->cdc[comp], 4);
      }
    if (cbpc == 2)
      for (int comp = 0; comp < 2; comp++)
        for (int blk = 0; blk < 4; blk++) {
          const int bx = blk & 1, by = blk >> 1;
          const uint8_t *cur = code->cnnz[comp];
          const int a = bx > 0 ? cur[by * 2] > 0 : (left ? left->cnnz[comp][by * 2 + 1] > 0 : un);
          const int bb = by > 0 ? cur[bx] > 0 : (top ? top->cnnz[comp][2 + bx] > 0 : un);
          write_block(c, 4, 85 + 16 + a + 2 * bb, code->cac[comp][blk] + 1, 15);
        }
  }
}

/* ---- parameter sets, slices, pictures (enc/mod.rs) ------------------------------------------------ */

/* (level_idc, MaxFS, MaxMBPS) of Table A-1 (enc/mod.rs:43-64) */
static const uint32_t LEVELS[19][3] = {
    {10, 99, 1485},       {11, 396, 3000},      {12, 396, 6000},     {13, 396, 11880},   {20, 396, 11880},
    {21, 792, 19800},     {22, 1620, 20250},    {30, 1620, 40500},   {31, 3600, 108000}, {32, 5120, 216000},
    {40, 8192, 245760},   {41, 8192, 245760},   {42, 8704, 522240},  {50, 22080, 589824}, {51, 36864, 983040},
    {52, 36864, 2073600}, {60, 139264, 4177920}, {61, 139264, 8355840}, {62, 139264, 16711680}};

struct iw_stream {
  iw_params p;
  int mb_w, mb_h, level, nslices;
  uint32_t frame_num, idr_id;
  int64_t idr_at;  /* display index of the last IDR picture */
  mbstate_t *state;
  bw_t out;
};

static double stream_fps(const iw_params *p) {
  return p->fps_num > 0 && p->fps_den > 0 ? (double)p->fps_num / (double)p->fps_den : 25.0;
}

int iw_level(int width, int height, int fps_num, int fps_den) {
  if (width < 1 || height < 1) return IW_EINVAL;
  const uint32_t fs = (uint32_t)(((width + 15) / 16) * ((height + 15) / 16));
  const double fps = fps_num > 0 && fps_den > 0 ? (double)fps_num / (double)fps_den : 25.0;
  double m = (double)fs * fps;
  uint64_t mbps = (uint64_t)m;
  if ((double)mbps < m) mbps++; /* ceil */
  for (int i = 0; i < 19; i++)
    if (LEVELS[i][1] >= fs && LEVELS[i][2] >= mbps) return (int)LEVELS[i][0];
  for (int i = 0; i < 19; i++)
    if (LEVELS[i][1] >= fs) return (int)LEVELS[i][0];
  return 62;
}

int iw_stream_create(iw_stream **out, const iw_params *p) {
  if (!out || !p || p->width < 16 || p->height < 16 || p->width > 16384 || p->height > 16384 || (p->width & 1) ||
      (p->height & 1) || p->nrefs < 1 || p->nrefs > 16 || p->slices < 1 || p->bframes < 0)
    return IW_EINVAL;
  if (!p->cabac) return IW_EUNSUPPORTED;
  iw_stream *s = (iw_stream *)calloc(1, sizeof *s);
  if (!s) return IW_ENOMEM;
  s->p = *p;
  s->mb_w = (p->width + 15) / 16;
  s->mb_h = (p->height + 15) / 16;
  s->nslices = p->slices < s->mb_h ? p->slices : s->mb_h;
  s->level = iw_level(p->width, p->height, p->fps_num, p->fps_den);
  s->state = (mbstate_t *)calloc((size_t)s->mb_w * s->mb_h, sizeof(mbstate_t));
  if (!s->state) {
    free(s);
    return IW_ENOMEM;
  }
  *out = s;
  return IW_OK;
}

void iw_stream_destroy(iw_stream *s) {
  if (!s) return;
  free(s->state);
  free(s->out.buf);
  free(s);
}

static void write_sps(iw_stream *s, bw_t *w) {
  const int high = s->p.cabac && s->p.transform8;
  if (high) {
    bw_put(w, 8, 100); /* High */
    bw_put(w, 8, 0);
  } else {
    bw_put(w, 8, 77); /* Main: constraint_set1 */
    bw_put(w, 8, 0x40);
  }
  bw_put(w, 8, (uint32_t)s->level);
  bw_ue(w, 0); /* seq_parameter_set_id */
  if (high) {
    bw_ue(w, 1); /* chroma_format_idc 4:2:0 */
    bw_ue(w, 0);
    bw_ue(w, 0);
    bw_flag(w, 0); /* qpprime_y_zero_transform_bypass_flag */
    bw_flag(w, 0); /* seq_scaling_matrix_present_flag */
  }
  bw_ue(w, 4); /* log2_max_frame_num_minus4: 8 bits */
  const int b = s->p.bframes > 0;
  /* the references kept: those P pictures use, and the two either side of B pictures */
  const uint32_t refs = (uint32_t)(b && s->p.nrefs < 2 ? 2 : s->p.nrefs);
  if (b) {
    bw_ue(w, 0); /* pic_order_cnt_type 0: counted in the slice headers */
    bw_ue(w, 4); /* log2_max_pic_order_cnt_lsb_minus4: 8 bits */
  } else {
    bw_ue(w, 2); /* pic_order_cnt_type 2: output order is decoding order */
  }
  bw_ue(w, refs); /* max_num_ref_frames */
  bw_flag(w, 0);
  bw_ue(w, (uint32_t)s->mb_w - 1);
  bw_ue(w, (uint32_t)s->mb_h - 1);
  bw_flag(w, 1); /* frame_mbs_only_flag */
  bw_flag(w, 1); /* direct_8x8_inference_flag */
  const uint32_t cr = (uint32_t)(s->mb_w * 16 - s->p.width) / 2, cb = (uint32_t)(s->mb_h * 16 - s->p.height) / 2;
  bw_flag(w, cr > 0 || cb > 0);
  if (cr > 0 || cb > 0) {
    bw_ue(w, 0);
    bw_ue(w, cr);
    bw_ue(w, 0);
    bw_ue(w, cb);
  }
  bw_flag(w, 1); /* vui_parameters_present_flag */
  bw_flag(w, 0);
  bw_flag(w, 0);
  bw_flag(w, 0);
  bw_flag(w, 0);
  const int64_t num = s->p.fps_num > 0 && s->p.fps_den > 0 ? s->p.fps_num : 25, den = s->p.fps_num > 0 && s->p.fps_den > 0 ? s->p.fps_den : 1;
  const int timing = den <= 0xFFFFFFFFll && 2 * num <= 0xFFFFFFFFll;
  bw_flag(w, timing);
  if (timing) {
    bw_put(w, 32, (uint32_t)den);     /* num_units_in_tick */
    bw_put(w, 32, (uint32_t)(2 * num)); /* time_scale */
    bw_flag(w, 1);                    /* fixed_frame_rate_flag */
  }
  bw_flag(w, 0);
  bw_flag(w, 0);
  bw_flag(w, 0);
  bw_flag(w, 1); /* bitstream_restriction_flag */
  bw_flag(w, 1);
  bw_ue(w, 0);
  bw_ue(w, 0);
  bw_ue(w, 15);
  bw_ue(w, 15);
  /* the P picture after B pictures is shown after them; B pictures are not kept */
  bw_ue(w, (uint32_t)b);                             /* max_num_reorder_frames */
  bw_ue(w, refs > 1u + (uint32_t)b ? refs : 1u + (uint32_t)b); /* max_dec_frame_buffering */
  bw_finish(w);
}

static void write_pps(iw_stream *s, bw_t *w) {
  bw_ue(w, 0);
  bw_ue(w, 0);
  bw_flag(w, s->p.cabac);
  bw_flag(w, 0);
  bw_ue(w, 0);
  bw_ue(w, 0); /* num_ref_idx_l0_default_active_minus1 */
  bw_ue(w, 0);
  bw_flag(w, 0);
  bw_put(w, 2, 0);
  bw_se(w, 0); /* pic_init_qp_minus26 */
  bw_se(w, 0);
  bw_se(w, 0); /* chroma_qp_index_offset */
  bw_flag(w, 0); /* deblocking_filter_control_present_flag */
  bw_flag(w, 0);
  bw_flag(w, 0);
  if (s->p.cabac && s->p.transform8) {
    bw_flag(w, 1); /* transform_8x8_mode_flag */
    bw_flag(w, 0);
    bw_se(w, 0);
  }
  bw_finish(w);
}

static int emit_nal(iw_stream *s, int ref_idc, int type, const bw_t *rbsp) {
  size_t need = 0;
  iw_put_nal(ref_idc, type, rbsp->buf, rbsp->len, NULL, 0, &need);
  while (s->out.cap < s->out.len + need) {
    const size_t nc = s->out.cap ? 2 * s->out.cap : 1 << 16;
    uint8_t *nb = (uint8_t *)realloc(s->out.buf, nc > s->out.len + need ? nc : s->out.len + need);
    if (!nb) return IW_ENOMEM;
    s->out.buf = nb;
    s->out.cap = nc > s->out.len + need ? nc : s->out.len + need;
  }
  size_t got = 0;
  iw_put_nal(ref_idc, type, rbsp->buf, rbsp->len, s->out.buf + s->out.len, need, &got);
  s->out.len += got;
  return IW_OK;
}

static int write_picture(iw_stream *s, int kind, int qp, int nref, int64_t display, const ia_mbcode *codes, size_t *len) {
  if (!s || !codes || !len || kind < 0 || kind > 2 || qp < 0 || qp > 51) return IW_EINVAL;
  const int idr = kind == 0, b = kind == 2, poc_type0 = s->p.bframes > 0;
  if (b && !poc_type0) return IW_EINVAL;
  if (kind == 1 && (nref < 1 || nref > s->p.nrefs)) return IW_EINVAL;
  if (idr) {
    s->frame_num = 0;
    s->idr_at = display;
    nref = 0;
  }
  /* B slices use one picture per list: the PPS's default of one active index each (enc/mod.rs:582-586) */
  if (b) nref = 1;
  if (poc_type0 && display < s->idr_at) return IW_EINVAL;
  const uint32_t poc = (uint32_t)(2 * (display - s->idr_at));
  s->out.len = 0;
  s->out.overflow = 0;
  int rc = IW_OK;
  if (idr) {
    bw_t w = {0};
    write_sps(s, &w);
    rc = w.overflow ? IW_ENOMEM : emit_nal(s, 3, 7, &w);
    free(w.buf);
    if (rc) return rc;
    bw_t q = {0};
    write_pps(s, &q);
    rc = q.overflow ? IW_ENOMEM : emit_nal(s, 3, 8, &q);
    free(q.buf);
    if (rc) return rc;
  }
  const int mb_w = s->mb_w, mb_h = s->mb_h, n = s->nslices;
  for (int i = 0; i < n && !rc; i++) {
    const int row0 = i * mb_h / n, row1 = (i + 1) * mb_h / n;
    cw_t c;
    memset(&c, 0, sizeof c);
    bw_t *w = &c.w;
    /* the slice header (enc/mod.rs:573-603) */
    bw_ue(w, (uint32_t)(row0 * mb_w));
    bw_ue(w, idr ? 7 : b ? 6 : 5); /* slice_type: I, B or P, all slices of the picture alike */
    bw_ue(w, 0);
    bw_put(w, 8, s->frame_num);
    if (idr) bw_ue(w, s->idr_id);
    if (poc_type0) bw_put(w, 8, poc & 255); /* pic_order_cnt_lsb */
    if (b) bw_flag(w, 1);                   /* direct_spatial_mv_pred_flag */
    if (!idr) {
      if (!b && nref > 1) {
        bw_flag(w, 1);
        bw_ue(w, (uint32_t)nref - 1);
      } else {
        bw_flag(w, 0);
      }
      bw_flag(w, 0);         /* ref_pic_list_modification_flag_l0 */
      if (b) bw_flag(w, 0);  /* ref_pic_list_modification_flag_l1 */
    }
    if (idr) {
      bw_flag(w, 0);
      bw_flag(w, 0);
    } else if (!b) {
      bw_flag(w, 0); /* adaptive_ref_pic_marking_mode_flag (B pictures are not references) */
    }
    if (!idr) bw_ue(w, 0); /* cabac_init_idc */
    bw_se(w, qp - 26);
    bw_align_ones(w);
    cw_init_contexts(&c, idr, qp);
    for (int mby = row0; mby < row1; mby++)
      for (int mbx = 0; mbx < mb_w; mbx++) {
        const int idx = mby * mb_w + mbx;
        const mbstate_t *left = mbx > 0 ? &s->state[idx - 1] : NULL;
        const mbstate_t *top = mby > row0 ? &s->state[idx - mb_w] : NULL;
        const ia_mbcode *code = &codes[idx];
        if (code->kind != K_I4 && code->kind != K_I8 && code->kind != K_I16 && code->kind != K_INTER) rc = IW_EINVAL;
        if (idr && code->kind == K_INTER) rc = IW_EINVAL;
        if (kind == 1 && code->kind == K_INTER && !code->skip)
          for (int k = 0; k < 4; k++)
            if (code->refs[k] >= nref) rc = IW_EINVAL;
        if (b && code->kind == K_INTER && !code->skip) {
          const int part = code->part & 3;
          for (int k = 0; k < (part == 0 ? 1 : part == 3 ? 4 : 2); k++)
            if (code->bdir[k] > 3 || (part != 0 && part != 3 && code->bdir[k] == 0)) rc = IW_EINVAL;
        }
        if (rc) break;
        mb_bins(&c, code, left, top, !idr, b, s->p.cabac && s->p.transform8, nref);
        cw_terminate(&c, mby + 1 == row1 && mbx + 1 == mb_w);
        finish_state(code, &s->state[idx], b);
      }
    bw_pad_zeros(w);
    if (!rc) rc = w->overflow ? IW_ENOMEM : emit_nal(s, idr ? 3 : b ? 0 : 2, idr ? 5 : 1, w);
    free(w->buf);
  }
  if (rc) {
    s->out.len = 0;
    return rc;
  }
  if (idr) s->idr_id = (s->idr_id + 1) % 65536;
  if (!b) s->frame_num = (s->frame_num + 1) % 256;
  *len = s->out.len;
  return IW_OK;
}

int iw_write_picture(iw_stream *s, int kind, int qp, int nref, const ia_mbcode *codes, size_t *len) {
  if (!s || kind == 2) return IW_EINVAL;
  return write_picture(s, kind, qp, nref, s->idr_at, codes, len);
}

int iw_write_picture_at(iw_stream *s, int kind, int qp, int nref, int64_t display, const ia_mbcode *codes,
                        size_t *len) {
  return write_picture(s, kind, qp, nref, display, codes, len);
}

int iw_read_picture(const iw_stream *s, uint8_t *out, size_t cap) {
  if (!s || !out) return IW_EINVAL;
  if (s->out.len > cap) return IW_ESPACE;
  memcpy(out, s->out.buf, s->out.len);
  return IW_OK;
}
