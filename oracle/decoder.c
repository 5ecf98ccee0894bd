/* decoder.c — the verifier: H.264 decoding of progressive 8-bit 4:2:0 I and P pictures coded with
 * CABAC, restated from nervosys/IronAV's decoder (crates/ironav-codec/src/h264/{mod,ps,bits,cabac,mb,
 * dpb}.rs; the loop filter through the oracle's restatement of h264/deblock.rs).
 *
 * TEST INFRASTRUCTURE ONLY (see decoder.h). Samples are held as uint16_t, as the reference holds them,
 * so the oracle's reconstruction primitives (intra prediction, interpolation, inverse transforms)
 * serve the decoder unchanged; the reference decoder's fixtures then pin them.
 */
#include "decoder.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "h264_dec_tables.h"
#include "oracle.h"

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int mini(int a, int b) { return a < b ? a : b; }
static inline int absi(int a) { return a < 0 ? -a : a; }
static inline int shl(int v, int s) { return (int)((uint32_t)v << s); }

enum { K_I4 = 1, K_I8 = 2, K_I16 = 3, K_PCM = 4, K_INTER = 5 };

/* ---- bits (h264/bits.rs) ------------------------------------------------------------------------ */

typedef struct {
  const uint8_t *p;
  size_t n;    /* bytes */
  size_t pos;  /* bits read */
} br_t;

static uint32_t br_bit(br_t *b) {
  if (b->pos >= b->n * 8) {
    b->pos++;
    return 0;
  }
  const uint32_t v = (b->p[b->pos >> 3] >> (7 - (b->pos & 7))) & 1;
  b->pos++;
  return v;
}

static uint32_t br_u(br_t *b, int n) {
  uint32_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 1) | br_bit(b);
  return v;
}

static uint32_t br_ue(br_t *b) {
  int lz = 0;
  while (br_bit(b) == 0) {
    if (++lz > 31) return 0xFFFFFFFFu;
  }
  return lz ? (1u << lz) - 1 + br_u(b, lz) : 0;
}

static int32_t br_se(br_t *b) {
  const uint32_t k = br_ue(b);
  return (k & 1) ? (int32_t)((k + 1) / 2) : -(int32_t)(k / 2);
}

static int br_more_rbsp(const br_t *b) {
  /* more data before the rbsp_stop_one_bit: some 1 bit after the current position other than the last */
  size_t last = b->n * 8;
  while (last > 0 && !((b->p[(last - 1) >> 3] >> (7 - ((last - 1) & 7))) & 1)) last--;
  return last > 0 && b->pos < last - 1;
}

/* emulation prevention removal (h264/bits.rs unescape): 0x000003 -> 0x0000 */
static size_t unescape(const uint8_t *in, size_t n, uint8_t *out) {
  size_t o = 0;
  int zeros = 0;
  for (size_t i = 0; i < n; i++) {
    if (zeros >= 2 && in[i] == 3) {
      zeros = 0;
      continue;
    }
    out[o++] = in[i];
    zeros = in[i] == 0 ? zeros + 1 : 0;
  }
  return o;
}

/* ---- parameter sets (h264/ps.rs) ---------------------------------------------------------------- */

typedef struct {
  int valid, profile, chroma_format_idc, bit_depth_luma, bit_depth_chroma, transform_bypass;
  int scaling_present;
  uint8_t sl4[6][16], sl8[6][64]; /* raster order */
  int log2_max_frame_num, poc_type, log2_max_poc_lsb, delta_pic_order_always_zero;
  int offset_for_non_ref, offset_top_to_bottom, ncycle, offset_for_ref[256];
  int max_num_ref_frames, gaps, mb_w, mb_h, frame_mbs_only, mbaff, direct8x8;
  int crop[4]; /* left, right, top, bottom in samples */
} sps_t;

typedef struct {
  int valid, sps_id, cabac, bottom_field_pic_order, num_slice_groups, num_ref_idx_default[2];
  int weighted_pred, weighted_bipred, pic_init_qp, chroma_qp_offset[2], deblock_ctrl, constrained_intra;
  int redundant_pic_cnt_present, transform8, scaling_present;
  uint8_t sl4[6][16], sl8[6][64];
} pps_t;

/* Scaling list syntax (7.3.2.1.1.1) into raster order: returns 1 when useDefaultScalingMatrixFlag. */
static int scaling_list(br_t *b, uint8_t *list, int size) {
  const uint8_t *scan = size == 16 ? OD_ZIGZAG4 : OD_ZIGZAG_DIRECT;
  int last = 8, next = 8;
  for (int j = 0; j < size; j++) {
    if (next != 0) {
      const int delta = br_se(b);
      next = (last + delta + 256) % 256;
      if (j == 0 && next == 0) return 1;
    }
    list[scan[j]] = (uint8_t)(next == 0 ? last : next);
    last = list[scan[j]];
  }
  return 0;
}

/* The scaling lists of an SPS or PPS with the fall-back rules (Table 7-2; rule A for the SPS with
 * flat fall-backs, rule B for the PPS with the SPS's lists). */
static void scaling_lists(br_t *b, int n8, uint8_t sl4[6][16], uint8_t sl8[6][64], const uint8_t fb4[6][16],
                          const uint8_t fb8[6][64]) {
  for (int i = 0; i < 6 + n8; i++) {
    const int present = (int)br_u(b, 1);
    if (i < 6) {
      if (!present) {
        if (i == 0 || i == 3) memcpy(sl4[i], fb4[i], 16);
        else memcpy(sl4[i], sl4[i - 1], 16);
      } else if (scaling_list(b, sl4[i], 16)) {
        memcpy(sl4[i], OD_DEFAULT_SCALING4[i < 3 ? 0 : 1], 16);
      }
    } else {
      const int k = i - 6;
      if (!present) {
        if (k < 2) memcpy(sl8[k], fb8[k], 64);
        else memcpy(sl8[k], sl8[k - 2], 64);
      } else if (scaling_list(b, sl8[k], 64)) {
        memcpy(sl8[k], OD_DEFAULT_SCALING8[k & 1], 64);
      }
    }
  }
}

static void flat_lists(uint8_t sl4[6][16], uint8_t sl8[6][64]) {
  memset(sl4, 16, 6 * 16);
  memset(sl8, 16, 6 * 64);
}

/* ---- the decoder ---------------------------------------------------------------------------------- */

typedef struct {
  int used, id, frame_num, frame_num_wrap, long_term, long_term_idx, is_ref, idr_seq, poc, order;
  int mb_w, mb_h, crop[4];
  uint8_t *y, *cb, *cr;       /* loop-filtered */
  uint16_t *y16, *cb16, *cr16; /* the same for interpolation */
  ia_mbinfo *mbs;
} frame_t;

typedef struct {
  int first_mb, slice_type, pps_id, frame_num, idr, idr_pic_id, poc_lsb, delta_poc_bottom, delta_poc[2];
  int nal_ref_idc, num_ref_idx_active;
  int nmods, mod_idc[32], mod_val[32];
  int luma_log2wd, chroma_log2wd;
  int lw[32], lo[32], cw[32][2], co[32][2];
  /* B slices: direct_spatial_mv_pred_flag, list 1's size, modifications and explicit weights */
  int direct_spatial, num_ref_idx_active1;
  int nmods1, mod_idc1[32], mod_val1[32];
  int lw1[32], lo1[32], cw1[32][2], co1[32][2];
  int no_output_prior, long_term_ref, adaptive, nmmco, mmco[32][3];
  int cabac_init_idc, qp, disable_deblock, alpha_offset, beta_offset;
} sh_t;

/* what the macroblock layer of later macroblocks and the loop filter read about a decoded one */
typedef struct {
  int slice;  /* 1-based slice number, 0 when not decoded */
  uint8_t kind, skip, cbp, t8, chroma_mode, cbf_dc, qp_delta_nz;
  uint8_t modes[16];
  uint8_t nz[16];      /* luma coefficients per 4x4 block */
  uint8_t cnz[2][4];
  uint8_t mvd[16][2];  /* |mvd| per 4x4 block, clipped */
  int8_t ref[4];
  int16_t mv[16][2];
  /* list 1 (B slices), and the direct prediction of B_Skip / B_Direct_16x16 and B_Direct_8x8 blocks */
  uint8_t mvd1[16][2];
  int8_t ref1[4];
  int16_t mv1[16][2];
  uint8_t bdirect, direct8;
  int qp;
} dmb_t;

#define OD_MAX_FRAMES 512

struct od_dec {
  sps_t sps[32];
  pps_t pps[256];
  char err[256];
  /* the picture being decoded */
  int have_cur, cur_sps, cur_pps;
  sh_t cur_hdr;
  int mb_w, mb_h;
  uint16_t *py, *pcb, *pcr; /* reconstruction before the loop filter */
  dmb_t *mb;
  ia_mbinfo *info;
  int nslices;
  or_slice_filter sfilter[4096];
  int cur_poc;
  /* order counts and references */
  int prev_poc_msb, prev_poc_lsb, prev_frame_num, prev_frame_num_offset, prev_ref_has_mmco5;
  int idr_seq, next_id;
  /* what the decoded syntax exercised: counts of Intra4x4 / Intra8x8 modes, Intra16x16 modes, chroma
     modes, P partitionings, sub-partitionings, P_Skip, I_PCM, 8x8-transform inter macroblocks,
     reference indexes above 0 */
  int64_t stats[OD_NSTATS];
  frame_t frames[OD_MAX_FRAMES];
  int nframes;
};

static int fail(od_dec *d, int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(d->err, sizeof d->err, fmt, ap);
  va_end(ap);
  return code;
}

od_dec *od_create(void) {
  od_dec *d = (od_dec *)calloc(1, sizeof *d);
  return d;
}

static void free_frame(frame_t *f) {
  free(f->y);
  free(f->cb);
  free(f->cr);
  free(f->y16);
  free(f->cb16);
  free(f->cr16);
  free(f->mbs);
  memset(f, 0, sizeof *f);
}

static void free_cur(od_dec *d) {
  free(d->py);
  free(d->pcb);
  free(d->pcr);
  free(d->mb);
  free(d->info);
  d->py = d->pcb = d->pcr = NULL;
  d->mb = NULL;
  d->info = NULL;
  d->have_cur = 0;
}

void od_destroy(od_dec *d) {
  if (!d) return;
  free_cur(d);
  for (int i = 0; i < d->nframes; i++) free_frame(&d->frames[i]);
  free(d);
}

const char *od_error(const od_dec *d) { return d ? d->err : "no decoder"; }

static int parse_sps(od_dec *d, br_t *b) {
  sps_t s;
  memset(&s, 0, sizeof s);
  s.profile = (int)br_u(b, 8);
  br_u(b, 16); /* constraint flags, level */
  const uint32_t id = br_ue(b);
  if (id > 31) return fail(d, OD_EBITSTREAM, "sps id %u", id);
  s.chroma_format_idc = 1;
  s.bit_depth_luma = s.bit_depth_chroma = 8;
  flat_lists(s.sl4, s.sl8);
  const int p = s.profile;
  if (p == 100 || p == 110 || p == 122 || p == 244 || p == 44 || p == 83 || p == 86 || p == 118 || p == 128 ||
      p == 138 || p == 139 || p == 134 || p == 135) {
    s.chroma_format_idc = (int)br_ue(b);
    if (s.chroma_format_idc == 3) br_u(b, 1);
    s.bit_depth_luma = 8 + (int)br_ue(b);
    s.bit_depth_chroma = 8 + (int)br_ue(b);
    s.transform_bypass = (int)br_u(b, 1);
    s.scaling_present = (int)br_u(b, 1);
    if (s.scaling_present) {
      uint8_t f4[6][16], f8[6][64];
      for (int i = 0; i < 6; i++) memcpy(f4[i], OD_DEFAULT_SCALING4[i < 3 ? 0 : 1], 16);
      for (int i = 0; i < 6; i++) memcpy(f8[i], OD_DEFAULT_SCALING8[i & 1], 64);
      scaling_lists(b, s.chroma_format_idc == 3 ? 6 : 2, s.sl4, s.sl8, (const uint8_t(*)[16])f4,
                    (const uint8_t(*)[64])f8);
    }
  }
  s.log2_max_frame_num = (int)br_ue(b) + 4;
  s.poc_type = (int)br_ue(b);
  if (s.poc_type == 0) {
    s.log2_max_poc_lsb = (int)br_ue(b) + 4;
  } else if (s.poc_type == 1) {
    s.delta_pic_order_always_zero = (int)br_u(b, 1);
    s.offset_for_non_ref = br_se(b);
    s.offset_top_to_bottom = br_se(b);
    s.ncycle = (int)br_ue(b);
    if (s.ncycle > 255) return fail(d, OD_EBITSTREAM, "poc cycle");
    for (int i = 0; i < s.ncycle; i++) s.offset_for_ref[i] = br_se(b);
  }
  s.max_num_ref_frames = (int)br_ue(b);
  s.gaps = (int)br_u(b, 1);
  s.mb_w = (int)br_ue(b) + 1;
  s.mb_h = (int)br_ue(b) + 1;
  s.frame_mbs_only = (int)br_u(b, 1);
  if (!s.frame_mbs_only) s.mbaff = (int)br_u(b, 1);
  s.direct8x8 = (int)br_u(b, 1);
  if (br_u(b, 1)) {
    for (int i = 0; i < 4; i++) s.crop[i] = (int)br_ue(b);
    /* CropUnitX = 2, CropUnitY = 2 for progressive 4:2:0 */
    for (int i = 0; i < 4; i++) s.crop[i] *= 2;
  }
  s.valid = 1;
  d->sps[id] = s;
  return OD_OK;
}

static int parse_pps(od_dec *d, br_t *b) {
  pps_t q;
  memset(&q, 0, sizeof q);
  const uint32_t id = br_ue(b);
  q.sps_id = (int)br_ue(b);
  if (id > 255 || q.sps_id > 31 || !d->sps[q.sps_id].valid) return fail(d, OD_EBITSTREAM, "pps %u", id);
  const sps_t *s = &d->sps[q.sps_id];
  q.cabac = (int)br_u(b, 1);
  q.bottom_field_pic_order = (int)br_u(b, 1);
  q.num_slice_groups = (int)br_ue(b) + 1;
  if (q.num_slice_groups != 1) return fail(d, OD_EUNSUPPORTED, "slice groups");
  q.num_ref_idx_default[0] = (int)br_ue(b) + 1;
  q.num_ref_idx_default[1] = (int)br_ue(b) + 1;
  q.weighted_pred = (int)br_u(b, 1);
  q.weighted_bipred = (int)br_u(b, 2);
  q.pic_init_qp = 26 + br_se(b);
  br_se(b); /* pic_init_qs */
  q.chroma_qp_offset[0] = q.chroma_qp_offset[1] = br_se(b);
  q.deblock_ctrl = (int)br_u(b, 1);
  q.constrained_intra = (int)br_u(b, 1);
  q.redundant_pic_cnt_present = (int)br_u(b, 1);
  memcpy(q.sl4, s->sl4, sizeof q.sl4);
  memcpy(q.sl8, s->sl8, sizeof q.sl8);
  if (br_more_rbsp(b)) {
    q.transform8 = (int)br_u(b, 1);
    q.scaling_present = (int)br_u(b, 1);
    if (q.scaling_present) {
      uint8_t f4[6][16], f8[6][64];
      if (s->scaling_present) {
        memcpy(f4, s->sl4, sizeof f4);
        memcpy(f8, s->sl8, sizeof f8);
      } else {
        for (int i = 0; i < 6; i++) memcpy(f4[i], OD_DEFAULT_SCALING4[i < 3 ? 0 : 1], 16);
        for (int i = 0; i < 6; i++) memcpy(f8[i], OD_DEFAULT_SCALING8[i & 1], 64);
      }
      scaling_lists(b, q.transform8 ? (s->chroma_format_idc == 3 ? 6 : 2) : 0, q.sl4, q.sl8,
                    (const uint8_t(*)[16])f4, (const uint8_t(*)[64])f8);
    }
    q.chroma_qp_offset[1] = br_se(b);
  }
  q.valid = 1;
  d->pps[id] = q;
  return OD_OK;
}

static int parse_slice_header(od_dec *d, br_t *b, int nal_type, int nal_ref_idc, sh_t *h) {
  memset(h, 0, sizeof *h);
  h->nal_ref_idc = nal_ref_idc;
  h->idr = nal_type == 5;
  h->first_mb = (int)br_ue(b);
  h->slice_type = (int)br_ue(b) % 5;
  if (h->slice_type > 2) return fail(d, OD_EUNSUPPORTED, "slice type %d", h->slice_type);
  h->pps_id = (int)br_ue(b);
  if (h->pps_id > 255 || !d->pps[h->pps_id].valid) return fail(d, OD_EBITSTREAM, "slice pps");
  const pps_t *q = &d->pps[h->pps_id];
  const sps_t *s = &d->sps[q->sps_id];
  if (!s->frame_mbs_only) return fail(d, OD_EUNSUPPORTED, "interlaced coding");
  if (s->chroma_format_idc != 1 || s->bit_depth_luma != 8 || s->bit_depth_chroma != 8 || s->transform_bypass)
    return fail(d, OD_EUNSUPPORTED, "only 8-bit 4:2:0");
  if (!q->cabac) return fail(d, OD_EUNSUPPORTED, "CAVLC");
  h->frame_num = (int)br_u(b, s->log2_max_frame_num);
  if (h->idr) h->idr_pic_id = (int)br_ue(b);
  if (s->poc_type == 0) {
    h->poc_lsb = (int)br_u(b, s->log2_max_poc_lsb);
    if (q->bottom_field_pic_order) h->delta_poc_bottom = br_se(b);
  }
  if (s->poc_type == 1 && !s->delta_pic_order_always_zero) {
    h->delta_poc[0] = br_se(b);
    if (q->bottom_field_pic_order) h->delta_poc[1] = br_se(b);
  }
  if (q->redundant_pic_cnt_present && br_ue(b) > 0) return fail(d, OD_EUNSUPPORTED, "redundant pictures");
  const int bslice = h->slice_type == 1;
  if (bslice) h->direct_spatial = (int)br_u(b, 1);
  h->num_ref_idx_active = q->num_ref_idx_default[0];
  h->num_ref_idx_active1 = q->num_ref_idx_default[1];
  if (h->slice_type != 2) {
    if (br_u(b, 1)) {
      h->num_ref_idx_active = (int)br_ue(b) + 1;
      if (bslice) h->num_ref_idx_active1 = (int)br_ue(b) + 1;
    }
    if (h->num_ref_idx_active > 32 || h->num_ref_idx_active1 > 32) return fail(d, OD_EBITSTREAM, "num_ref_idx");
    /* ref_pic_list_modification (7.3.3.1): list 0, and list 1 of a B slice */
    for (int l = 0; l < (bslice ? 2 : 1); l++) {
      int *n = l ? &h->nmods1 : &h->nmods, *idcs = l ? h->mod_idc1 : h->mod_idc, *vals = l ? h->mod_val1 : h->mod_val;
      if (br_u(b, 1)) {
        for (;;) {
          const uint32_t idc = br_ue(b);
          if (idc == 3) break;
          if (idc > 2 || *n >= 32) return fail(d, OD_EBITSTREAM, "ref list modification");
          idcs[*n] = (int)idc;
          vals[*n] = (int)br_ue(b);
          (*n)++;
        }
      }
    }
    /* pred_weight_table (7.3.3.2): P slices with weighted_pred_flag, B slices with weighted_bipred_idc 1 */
    for (int i = 0; i < 32; i++) {
      h->lw[i] = h->lw1[i] = 1;
      h->lo[i] = h->lo1[i] = 0;
    }
    if ((q->weighted_pred && !bslice) || (q->weighted_bipred == 1 && bslice)) {
      h->luma_log2wd = (int)br_ue(b);
      h->chroma_log2wd = (int)br_ue(b);
      for (int l = 0; l < (bslice ? 2 : 1); l++) {
        int *lw = l ? h->lw1 : h->lw, *lo = l ? h->lo1 : h->lo;
        int(*cw)[2] = l ? h->cw1 : h->cw, (*co)[2] = l ? h->co1 : h->co;
        for (int i = 0; i < (l ? h->num_ref_idx_active1 : h->num_ref_idx_active); i++) {
          lw[i] = 1 << h->luma_log2wd;
          lo[i] = 0;
          if (br_u(b, 1)) {
            lw[i] = br_se(b);
            lo[i] = br_se(b);
          }
          cw[i][0] = cw[i][1] = 1 << h->chroma_log2wd;
          co[i][0] = co[i][1] = 0;
          if (br_u(b, 1))
            for (int c = 0; c < 2; c++) {
              cw[i][c] = br_se(b);
              co[i][c] = br_se(b);
            }
        }
      }
    }
  }
  if (nal_ref_idc) {
    /* dec_ref_pic_marking (7.3.3.3) */
    if (h->idr) {
      h->no_output_prior = (int)br_u(b, 1);
      h->long_term_ref = (int)br_u(b, 1);
    } else if ((h->adaptive = (int)br_u(b, 1))) {
      for (;;) {
        const uint32_t op = br_ue(b);
        if (op == 0) break;
        if (op > 6 || h->nmmco >= 32) return fail(d, OD_EBITSTREAM, "mmco");
        int *m = h->mmco[h->nmmco++];
        m[0] = (int)op;
        m[1] = m[2] = 0;
        if (op == 1 || op == 3) m[1] = (int)br_ue(b);
        if (op == 2) m[1] = (int)br_ue(b);
        if (op == 3 || op == 6) m[2] = (int)br_ue(b);
        if (op == 4) m[1] = (int)br_ue(b);
      }
    }
  }
  if (q->cabac && h->slice_type != 2) h->cabac_init_idc = (int)br_ue(b);
  h->qp = q->pic_init_qp + br_se(b);
  if (q->deblock_ctrl) {
    h->disable_deblock = (int)br_ue(b);
    if (h->disable_deblock != 1) {
      h->alpha_offset = br_se(b) * 2;
      h->beta_offset = br_se(b) * 2;
    }
  }
  if (h->qp < 0 || h->qp > 51) return fail(d, OD_EBITSTREAM, "slice qp %d", h->qp);
  return OD_OK;
}

/* ---- order counts, reference lists and marking (h264/dpb.rs, 8.2.1, 8.2.4, 8.2.5) ---------------- */

static int compute_poc(od_dec *d, const sps_t *s, const sh_t *h) {
  if (s->poc_type == 0) {
    if (h->idr) {
      d->prev_poc_msb = 0;
      d->prev_poc_lsb = 0;
    }
    const int max_lsb = 1 << s->log2_max_poc_lsb;
    int msb;
    if (h->poc_lsb < d->prev_poc_lsb && d->prev_poc_lsb - h->poc_lsb >= max_lsb / 2) msb = d->prev_poc_msb + max_lsb;
    else if (h->poc_lsb > d->prev_poc_lsb && h->poc_lsb - d->prev_poc_lsb > max_lsb / 2) msb = d->prev_poc_msb - max_lsb;
    else msb = d->prev_poc_msb;
    const int top = msb + h->poc_lsb, bot = top + h->delta_poc_bottom;
    if (h->nal_ref_idc) {
      d->prev_poc_msb = msb;
      d->prev_poc_lsb = h->poc_lsb;
    }
    return mini(top, bot);
  }
  const int max_fn = 1 << s->log2_max_frame_num;
  int fno;
  if (h->idr) fno = 0;
  else if (d->prev_frame_num > h->frame_num) fno = d->prev_frame_num_offset + max_fn;
  else fno = d->prev_frame_num_offset;
  int poc;
  if (s->poc_type == 2) {
    poc = h->idr ? 0 : (h->nal_ref_idc ? 2 * (fno + h->frame_num) : 2 * (fno + h->frame_num) - 1);
  } else {
    int abs_fn = s->ncycle ? fno + h->frame_num : 0;
    if (!h->nal_ref_idc && abs_fn > 0) abs_fn--;
    int expected = 0;
    if (abs_fn > 0) {
      int delta_cycle = 0;
      for (int i = 0; i < s->ncycle; i++) delta_cycle += s->offset_for_ref[i];
      const int cnt = (abs_fn - 1) / s->ncycle, in = (abs_fn - 1) % s->ncycle;
      expected = cnt * delta_cycle;
      for (int i = 0; i <= in; i++) expected += s->offset_for_ref[i];
    }
    if (!h->nal_ref_idc) expected += s->offset_for_non_ref;
    const int top = expected + h->delta_poc[0];
    const int bot = top + s->offset_top_to_bottom + h->delta_poc[1];
    poc = mini(top, bot);
  }
  d->prev_frame_num_offset = fno;
  d->prev_frame_num = h->frame_num;
  return poc;
}

static int num_refs(const od_dec *d) {
  int n = 0;
  for (int i = 0; i < d->nframes; i++) n += d->frames[i].is_ref != 0;
  return n;
}

/* The modifications of one list (8.2.4.3.1 and 8.2.4.3.2) on its initial entries (indexes into
 * d->frames, -1 past the initial list), then the check that na entries name pictures. */
static int apply_mods(od_dec *d, const sps_t *s, const sh_t *h, int *list, int na, int nmods, const int *idcs,
                      const int *vals) {
  const int max_fn = 1 << s->log2_max_frame_num;
  int pred = h->frame_num, idx = 0;
  for (int k = 0; k < nmods; k++) {
    int target = -1;
    if (idcs[k] < 2) {
      const int diff = vals[k] + 1;
      int no_wrap = idcs[k] == 0 ? pred - diff : pred + diff;
      if (no_wrap < 0) no_wrap += max_fn;
      if (no_wrap >= max_fn) no_wrap -= max_fn;
      pred = no_wrap;
      const int pic_num = no_wrap > h->frame_num ? no_wrap - max_fn : no_wrap;
      for (int i = 0; i < d->nframes; i++)
        if (d->frames[i].is_ref && !d->frames[i].long_term && d->frames[i].frame_num_wrap == pic_num) target = i;
    } else {
      for (int i = 0; i < d->nframes; i++)
        if (d->frames[i].is_ref && d->frames[i].long_term && d->frames[i].long_term_idx == vals[k]) target = i;
    }
    if (target < 0) return fail(d, OD_EBITSTREAM, "modification names no reference");
    for (int i = na; i > idx; i--) list[i] = list[i - 1];
    list[idx++] = target;
    int o = idx;
    for (int i = idx; i <= na; i++)
      if (list[i] != target) list[o++] = list[i];
  }
  for (int i = 0; i < na; i++)
    if (list[i] < 0) return fail(d, OD_EBITSTREAM, "reference list shorter than num_ref_idx_active");
  return OD_OK;
}

/* Short-term references (with FrameNumWrap set) and long-term ones, unordered. */
static void collect_refs(od_dec *d, const sps_t *s, const sh_t *h, int *st, int *ns, int *lt, int *nl) {
  const int max_fn = 1 << s->log2_max_frame_num;
  *ns = *nl = 0;
  for (int i = 0; i < d->nframes; i++) {
    frame_t *f = &d->frames[i];
    if (!f->is_ref) continue;
    if (f->long_term) {
      lt[(*nl)++] = i;
    } else {
      f->frame_num_wrap = f->frame_num > h->frame_num ? f->frame_num - max_fn : f->frame_num;
      st[(*ns)++] = i;
    }
  }
}

/* Sorts frame indexes by key(frame) ascending (descending when desc). */
static void sort_frames(od_dec *d, int *v, int n, int key /* 0 FrameNumWrap, 1 LongTermPicNum, 2 order count */,
                        int desc) {
  for (int a = 0; a < n; a++)
    for (int b2 = a + 1; b2 < n; b2++) {
      const frame_t *x = &d->frames[v[a]], *y = &d->frames[v[b2]];
      const int kx = key == 0 ? x->frame_num_wrap : key == 1 ? x->long_term_idx : x->poc;
      const int ky = key == 0 ? y->frame_num_wrap : key == 1 ? y->long_term_idx : y->poc;
      if (desc ? ky > kx : ky < kx) {
        const int t = v[a];
        v[a] = v[b2];
        v[b2] = t;
      }
    }
}

/* The list-0 references of a P slice (8.2.4.2.1 and 8.2.4.3): short-term by descending PicNum, then
 * long-term by ascending LongTermPicNum, then the modifications; indexes into d->frames. */
static int ref_list(od_dec *d, const sps_t *s, const sh_t *h, int *list /* 33 entries */) {
  int st[OD_MAX_FRAMES], lt[OD_MAX_FRAMES], ns, nl;
  collect_refs(d, s, h, st, &ns, lt, &nl);
  sort_frames(d, st, ns, 0, 1);
  sort_frames(d, lt, nl, 1, 0);
  int init[64], n = 0;
  for (int i = 0; i < ns && n < 64; i++) init[n++] = st[i];
  for (int i = 0; i < nl && n < 64; i++) init[n++] = lt[i];
  const int na = h->num_ref_idx_active;
  if (n == 0) return fail(d, OD_EBITSTREAM, "P slice without references");
  for (int i = 0; i <= na && i < 33; i++) list[i] = i < n && i < na ? init[i] : -1;
  return apply_mods(d, s, h, list, na, h->nmods, h->mod_idc, h->mod_val);
}

/* The two lists of a B slice (8.2.4.2.3, frames): list 0 the short-term references shown before the
 * picture by descending order count, then those after it by ascending count, then the long-term
 * ones; list 1 the same with before and after exchanged, its first two entries swapped when it has
 * more than one and equals list 0; then each list's modifications. */
static int ref_lists_b(od_dec *d, const sps_t *s, const sh_t *h, int *list0, int *list1) {
  int st[OD_MAX_FRAMES], lt[OD_MAX_FRAMES], ns, nl;
  collect_refs(d, s, h, st, &ns, lt, &nl);
  sort_frames(d, lt, nl, 1, 0);
  int before[OD_MAX_FRAMES], after[OD_MAX_FRAMES], nb = 0, na_ = 0;
  for (int i = 0; i < ns; i++) {
    if (d->frames[st[i]].poc <= d->cur_poc) before[nb++] = st[i];
    else after[na_++] = st[i];
  }
  sort_frames(d, before, nb, 2, 1);
  sort_frames(d, after, na_, 2, 0);
  int init[2][64], n[2] = {0, 0};
  for (int l = 0; l < 2; l++) {
    const int *first = l ? after : before, *second = l ? before : after;
    const int nf = l ? na_ : nb, nsd = l ? nb : na_;
    for (int i = 0; i < nf && n[l] < 64; i++) init[l][n[l]++] = first[i];
    for (int i = 0; i < nsd && n[l] < 64; i++) init[l][n[l]++] = second[i];
    for (int i = 0; i < nl && n[l] < 64; i++) init[l][n[l]++] = lt[i];
  }
  if (n[0] == 0) return fail(d, OD_EBITSTREAM, "B slice without references");
  if (n[1] > 1) {
    int same = n[0] == n[1];
    for (int i = 0; same && i < n[1]; i++) same = init[0][i] == init[1][i];
    if (same) {
      const int t = init[1][0];
      init[1][0] = init[1][1];
      init[1][1] = t;
    }
  }
  const int nact[2] = {h->num_ref_idx_active, h->num_ref_idx_active1};
  for (int l = 0; l < 2; l++) {
    int *list = l ? list1 : list0;
    for (int i = 0; i <= nact[l] && i < 33; i++) list[i] = i < n[l] && i < nact[l] ? init[l][i] : -1;
  }
  int rc = apply_mods(d, s, h, list0, nact[0], h->nmods, h->mod_idc, h->mod_val);
  if (!rc) rc = apply_mods(d, s, h, list1, nact[1], h->nmods1, h->mod_idc1, h->mod_val1);
  return rc;
}

/* Marking after the picture (8.2.5): IDR, sliding window, or the adaptive operations. */
static void mark(od_dec *d, const sps_t *s, const sh_t *h, frame_t *cur) {
  const int max_fn = 1 << s->log2_max_frame_num;
  if (!h->nal_ref_idc) return;
  if (h->idr) {
    for (int i = 0; i < d->nframes; i++)
      if (&d->frames[i] != cur) d->frames[i].is_ref = 0;
    cur->is_ref = 1;
    cur->long_term = h->long_term_ref;
    cur->long_term_idx = 0;
    return;
  }
  if (h->adaptive) {
    for (int k = 0; k < h->nmmco; k++) {
      const int *m = h->mmco[k];
      if (m[0] == 1 || m[0] == 3) {
        int pic_num = h->frame_num - (m[1] + 1);
        for (int i = 0; i < d->nframes; i++) {
          frame_t *f = &d->frames[i];
          if (f == cur || !f->is_ref || f->long_term) continue;
          const int wrap = f->frame_num > h->frame_num ? f->frame_num - max_fn : f->frame_num;
          if (wrap == pic_num) {
            if (m[0] == 1) f->is_ref = 0;
            else {
              f->long_term = 1;
              f->long_term_idx = m[2];
            }
          }
        }
      } else if (m[0] == 2) {
        for (int i = 0; i < d->nframes; i++)
          if (d->frames[i].is_ref && d->frames[i].long_term && d->frames[i].long_term_idx == m[1]) d->frames[i].is_ref = 0;
      } else if (m[0] == 4) {
        for (int i = 0; i < d->nframes; i++)
          if (d->frames[i].is_ref && d->frames[i].long_term && d->frames[i].long_term_idx > m[1] - 1)
            d->frames[i].is_ref = 0;
      } else if (m[0] == 5) {
        for (int i = 0; i < d->nframes; i++)
          if (&d->frames[i] != cur) d->frames[i].is_ref = 0;
      } else if (m[0] == 6) {
        cur->long_term = 1;
        cur->long_term_idx = m[2];
      }
    }
    cur->is_ref = 1;
    return;
  }
  /* sliding window (8.2.5.3) */
  cur->is_ref = 1;
  while (num_refs(d) > (s->max_num_ref_frames > 0 ? s->max_num_ref_frames : 1)) {
    int victim = -1, best = 0x7fffffff;
    for (int i = 0; i < d->nframes; i++) {
      frame_t *f = &d->frames[i];
      if (f == cur || !f->is_ref || f->long_term) continue;
      const int wrap = f->frame_num > h->frame_num ? f->frame_num - max_fn : f->frame_num;
      if (wrap < best) {
        best = wrap;
        victim = i;
      }
    }
    if (victim < 0) break;
    d->frames[victim].is_ref = 0;
  }
}

/* ---- the CABAC decoding engine (h264/cabac.rs, 9.3.1.2 and 9.3.3.2) ------------------------------ */

typedef struct {
  br_t *b;
  uint32_t range, offset;
  uint8_t st[1024]; /* pStateIdx << 1 | valMPS */
} cabac_t;

static void cabac_init_contexts(cabac_t *c, int islice, int idc, int qp) {
  qp = clampi(qp, 0, 51);
  for (int i = 0; i < 1024; i++) {
    const int m = islice ? OD_CABAC_INIT_I[i][0] : OD_CABAC_INIT_PB[idc][i][0];
    const int n = islice ? OD_CABAC_INIT_I[i][1] : OD_CABAC_INIT_PB[idc][i][1];
    const int pre = clampi(((m * qp) >> 4) + n, 1, 126);
    c->st[i] = pre <= 63 ? (uint8_t)((63 - pre) << 1) : (uint8_t)(((pre - 64) << 1) | 1);
  }
}

static void cabac_init_engine(cabac_t *c) {
  c->range = 510;
  c->offset = br_u(c->b, 9);
}

static int decision(cabac_t *c, int ctx) {
  const int s = c->st[ctx], p = s >> 1, mps = s & 1;
  const uint32_t lps = OD_LPS_RANGE[p][(c->range >> 6) & 3];
  int bin;
  c->range -= lps;
  if (c->offset >= c->range) {
    bin = !mps;
    c->offset -= c->range;
    c->range = lps;
    c->st[ctx] = (uint8_t)((OD_TRANS_LPS[p] << 1) | (p == 0 ? !mps : mps));
  } else {
    bin = mps;
    c->st[ctx] = (uint8_t)(((p < 62 ? p + 1 : p) << 1) | mps);
  }
  while (c->range < 256) {
    c->range <<= 1;
    c->offset = (c->offset << 1) | br_bit(c->b);
  }
  return bin;
}

static int bypass(cabac_t *c) {
  c->offset = (c->offset << 1) | br_bit(c->b);
  if (c->offset >= c->range) {
    c->offset -= c->range;
    return 1;
  }
  return 0;
}

static int terminate(cabac_t *c) {
  c->range -= 2;
  if (c->offset >= c->range) return 1;
  while (c->range < 256) {
    c->range <<= 1;
    c->offset = (c->offset << 1) | br_bit(c->b);
  }
  return 0;
}

static uint32_t exp_golomb_bypass(cabac_t *c, int k) {
  uint32_t v = 0;
  while (bypass(c)) {
    v += 1u << k;
    if (++k > 30) return v;
  }
  while (k--) v += (uint32_t)bypass(c) << k;
  return v;
}

/* ---- the slice decoder (h264/mb.rs) --------------------------------------------------------------- */

typedef struct {
  od_dec *d;
  const sps_t *s;
  const pps_t *q;
  const sh_t *h;
  cabac_t c;
  int slice_id, mb_w, mb_h, mbx, mby, addr, qp, prev_qp_delta_nz, prev_addr;
  int refs[33]; /* frame indexes of list 0 (one spare entry for the modification shifts) */
  int refs1[33]; /* list 1 of a B slice */
  int bslice;
  dmb_t *cur;
  /* the macroblock being decoded */
  int32_t coef[16][16];  /* luma 4x4 blocks, raster block order, raster positions (dequantised later) */
  int32_t coef8[4][64];
  int32_t dc[16];        /* Intra16x16 DC levels, raster 4x4 */
  int32_t cdc[2][4];
  int32_t cac[2][4][16];
  int i16_mode;
  uint32_t done;  /* 4x4 blocks whose list 0 motion is known, for neighbour availability inside the macroblock */
  uint32_t done1; /* the same for list 1 */
} sdec_t;

static dmb_t *mb_at(sdec_t *t, int mbx, int mby) {
  if (mbx < 0 || mby < 0 || mbx >= t->mb_w || mby >= t->mb_h) return NULL;
  dmb_t *m = &t->d->mb[mby * t->mb_w + mbx];
  return m->slice == t->slice_id ? m : NULL;
}

static dmb_t *left_mb(sdec_t *t) { return mb_at(t, t->mbx - 1, t->mby); }
static dmb_t *top_mb(sdec_t *t) { return mb_at(t, t->mbx, t->mby - 1); }

/* The 4x4 block at (x, y) relative to the macroblock (x, y in -1..4): its macroblock (NULL when not
 * available) and its position there. Blocks of the current macroblock are available once `done`. */
static dmb_t *nb4(sdec_t *t, int x, int y, int *bx, int *by, int inside_done) {
  int dx = 0, dy = 0;
  if (x < 0) { dx = -1; x += 4; }
  else if (x > 3) { dx = 1; x -= 4; }
  if (y < 0) { dy = -1; y += 4; }
  else if (y > 3) return NULL;
  if (dx == 1 && dy == 0) return NULL;
  *bx = x;
  *by = y;
  if (dx == 0 && dy == 0) {
    /* inside_done: 1 blocks of known list 0 motion, 2 of known list 1 motion */
    if (inside_done && !(((inside_done == 2 ? t->done1 : t->done) >> (y * 4 + x)) & 1)) return NULL;
    return t->cur;
  }
  return mb_at(t, t->mbx + dx, t->mby + dy);
}

/* -- syntax elements (9.3.2, 9.3.3.1) -- */

static int read_mb_skip(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t);
  return decision(&t->c, (t->bslice ? 24 : 11) + (a && !a->skip) + (b && !b->skip));
}

/* I-slice mb_type bins after the first (9.3.2.5, Table 9-36), ctxIdx base 3 (I slice) or 17 (prefix in
 * P slices); returns 0..25 as the I mb_type */
static int read_i_mb_type_at(sdec_t *t, int islice, int prefix_base) {
  int first;
  if (islice) {
    const dmb_t *a = left_mb(t), *b = top_mb(t);
    const int ca = a && a->kind != K_I4 && a->kind != K_I8, cb = b && b->kind != K_I4 && b->kind != K_I8;
    first = decision(&t->c, 3 + ca + cb);
  } else {
    first = decision(&t->c, prefix_base);
  }
  if (!first) return 0;
  if (terminate(&t->c)) return 25;
  const int base = islice ? 3 : prefix_base;
  const int b_cbpl = islice ? base + 3 : base + 1;
  int mt = 1;
  const int luma = decision(&t->c, b_cbpl);
  const int chroma = decision(&t->c, islice ? base + 4 : base + 2);
  int cc = 0;
  if (chroma) cc = 1 + decision(&t->c, islice ? base + 5 : base + 2);
  const int m1 = decision(&t->c, islice ? base + 6 : base + 3);
  const int m0 = decision(&t->c, islice ? base + 7 : base + 3);
  mt += (m1 << 1 | m0) + 4 * cc + 12 * luma;
  return mt;
}

static int read_i_mb_type(sdec_t *t, int islice) { return read_i_mb_type_at(t, islice, 17); }

/* P mb_type (Table 9-37): 0..3 the P types (16x16, 16x8, 8x16, 8x8), 5 + I mb_type for intra */
static int read_p_mb_type(sdec_t *t) {
  if (decision(&t->c, 14)) return 5 + read_i_mb_type(t, 0);
  const int b1 = decision(&t->c, 15);
  const int b2 = decision(&t->c, b1 ? 17 : 16);
  if (!b1) return b2 ? 3 : 0;
  return b2 ? 1 : 2;
}

static int read_sub_mb_type(sdec_t *t) {
  if (decision(&t->c, 21)) return 0;
  if (!decision(&t->c, 22)) return 1;
  return decision(&t->c, 23) ? 2 : 3;
}

static int read_t8_flag(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t);
  return decision(&t->c, 399 + (a && a->t8) + (b && b->t8));
}

static int read_chroma_mode(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t);
  const int ca = a && a->kind != K_INTER && a->kind != K_PCM && a->chroma_mode != 0;
  const int cb = b && b->kind != K_INTER && b->kind != K_PCM && b->chroma_mode != 0;
  if (!decision(&t->c, 64 + ca + cb)) return 0;
  if (!decision(&t->c, 67)) return 1;
  return decision(&t->c, 67) ? 3 : 2;
}

/* coded_block_pattern (9.3.2.6, 9.3.3.1.1.4) */
static int read_cbp(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t);
  int luma = 0;
  for (int b8 = 0; b8 < 4; b8++) {
    int ca, cb;
    if (b8 & 1) ca = !((luma >> (b8 - 1)) & 1);
    else ca = a && (a->kind == K_PCM ? 0 : (a->skip || !((a->cbp >> (b8 + 1)) & 1)));
    if (b8 >= 2) cb = !((luma >> (b8 - 2)) & 1);
    else cb = b && (b->kind == K_PCM ? 0 : (b->skip || !((b->cbp >> (b8 + 2)) & 1)));
    luma |= decision(&t->c, 73 + ca + 2 * cb) << b8;
  }
  int ca = a && (a->kind == K_PCM || (!a->skip && (a->cbp >> 4) != 0));
  int cb = b && (b->kind == K_PCM || (!b->skip && (b->cbp >> 4) != 0));
  int chroma = 0;
  if (decision(&t->c, 77 + ca + 2 * cb)) {
    ca = a && (a->kind == K_PCM || (!a->skip && (a->cbp >> 4) == 2));
    cb = b && (b->kind == K_PCM || (!b->skip && (b->cbp >> 4) == 2));
    chroma = 1 + decision(&t->c, 81 + ca + 2 * cb);
  }
  return luma | chroma << 4;
}

static int read_qp_delta(sdec_t *t) {
  int k = 0;
  if (decision(&t->c, 60 + t->prev_qp_delta_nz)) {
    k = 1;
    int ctx = 62;
    while (decision(&t->c, ctx)) {
      ctx = 63;
      if (++k > 104) break;
    }
  }
  return (k & 1) ? (k + 1) / 2 : -(k / 2);
}

static int read_ref_idx(sdec_t *t, int bx, int by) {
  int nx, ny;
  const dmb_t *a = nb4(t, bx - 1, by, &nx, &ny, 0);
  const int ca = a && a->kind == K_INTER && !a->skip && a->ref[(ny >> 1) * 2 + (nx >> 1)] > 0;
  const dmb_t *b = nb4(t, bx, by - 1, &nx, &ny, 0);
  const int cb = b && b->kind == K_INTER && !b->skip && b->ref[(ny >> 1) * 2 + (nx >> 1)] > 0;
  int r = 0, ctx = 54 + ca + 2 * cb;
  while (decision(&t->c, ctx)) {
    r++;
    ctx = r == 1 ? 58 : 59;
    if (r > 32) break;
  }
  return r;
}

static int read_mvd(sdec_t *t, int bx, int by, int comp) {
  int nx, ny;
  const dmb_t *a = nb4(t, bx - 1, by, &nx, &ny, 0);
  int sum = a ? a->mvd[ny * 4 + nx][comp] : 0;
  const dmb_t *b = nb4(t, bx, by - 1, &nx, &ny, 0);
  sum += b ? b->mvd[ny * 4 + nx][comp] : 0;
  const int base = comp ? 47 : 40;
  const int inc = sum < 3 ? 0 : (sum > 32 ? 2 : 1);
  if (!decision(&t->c, base + inc)) return 0;
  int v = 1, ctx = base + 3;
  while (v < 9 && decision(&t->c, ctx)) {
    v++;
    if (ctx < base + 6) ctx++;
  }
  if (v >= 9) v += (int)exp_golomb_bypass(&t->c, 3);
  return bypass(&t->c) ? -v : v;
}

/* residual_block_cabac (7.3.5.3.3, 9.3.3.1.1.9, 9.3.3.1.3): cat 0..5, `cbf_inc` the coded_block_flag
 * increment (negative: no flag, cat 5); levels in scan order into `out` (max entries) */
static int read_block(sdec_t *t, int cat, int cbf_inc, int max, int32_t *out) {
  static const int cat_off[5] = {0, 4, 8, 12, 16};
  for (int i = 0; i < max; i++) out[i] = 0;
  if (cbf_inc >= 0 && !decision(&t->c, 85 + cat_off[cat] + cbf_inc)) return 0;
  int sig[64], n = 0, last = -1;
  for (int i = 0; i < max - 1; i++) {
    int si, li;
    if (cat == 3) si = li = mini(i, 2);
    else if (cat == 5) { si = OD_SIG_OFF_8X8[0][i]; li = OD_LAST_OFF_8X8[i]; }
    else si = li = i;
    sig[i] = decision(&t->c, OD_SIG_OFF[0][cat] + si);
    if (sig[i]) {
      n++;
      if (decision(&t->c, OD_LAST_OFF[0][cat] + li)) {
        last = i;
        break;
      }
    }
  }
  if (last < 0) {
    last = max - 1;
    sig[last] = 1;
    n++;
  }
  const int base = OD_ABS_OFF[cat];
  int eq1 = 0, gt1 = 0;
  for (int i = last; i >= 0; i--) {
    if (!sig[i]) continue;
    int a;
    if (!decision(&t->c, base + (gt1 ? 0 : mini(4, 1 + eq1)))) {
      a = 1;
      eq1++;
    } else {
      const int ctx = base + 5 + mini(gt1, 4 - (cat == 3));
      int k = 1;
      while (k < 14 && decision(&t->c, ctx)) k++;
      a = k + 1;
      if (k >= 14) a += (int)exp_golomb_bypass(&t->c, 0);
      gt1++;
    }
    out[i] = bypass(&t->c) ? -a : a;
  }
  return n;
}

/* coded_block_flag increment of a luma 4x4 block (cat 1/2) at (bx, by) */
static int cbf_luma4(sdec_t *t, int bx, int by, int intra) {
  int cond[2];
  for (int k = 0; k < 2; k++) {
    int nx, ny;
    const dmb_t *n = nb4(t, bx - (k == 0), by - (k == 1), &nx, &ny, 0);
    if (!n) cond[k] = intra;
    else if (n->kind == K_PCM) cond[k] = 1;
    else if (n->skip) cond[k] = 0;
    else if (!((n->cbp >> ((ny >> 1) * 2 + (nx >> 1))) & 1)) cond[k] = 0;
    else cond[k] = n->nz[ny * 4 + nx] != 0; /* an 8x8 transform block with coefficients counts as coded */
  }
  return cond[0] + 2 * cond[1];
}

static int cbf_dc(sdec_t *t, int bit, int intra, int chroma) {
  const dmb_t *n[2] = {left_mb(t), top_mb(t)};
  int cond[2];
  for (int k = 0; k < 2; k++) {
    if (!n[k]) cond[k] = intra;
    else if (n[k]->kind == K_PCM) cond[k] = 1;
    else if (n[k]->skip) cond[k] = 0;
    else if (!chroma && n[k]->kind != K_I16) cond[k] = 0;
    else if (chroma && (n[k]->cbp >> 4) == 0) cond[k] = 0;
    else cond[k] = (n[k]->cbf_dc >> bit) & 1;
  }
  return cond[0] + 2 * cond[1];
}

static int cbf_chroma_ac(sdec_t *t, int comp, int bx, int by, int intra) {
  int cond[2];
  for (int k = 0; k < 2; k++) {
    int x = bx - (k == 0), y = by - (k == 1);
    const dmb_t *n;
    if (x >= 0 && y >= 0) n = t->cur;
    else if (x < 0) { n = left_mb(t); x += 2; }
    else { n = top_mb(t); y += 2; }
    if (!n) cond[k] = intra;
    else if (n->kind == K_PCM) cond[k] = 1;
    else if (n->skip || (n->cbp >> 4) != 2) cond[k] = 0;
    else cond[k] = n->cnz[comp][y * 2 + x] != 0;
  }
  return cond[0] + 2 * cond[1];
}

/* -- motion vector prediction (8.4.1.3, h264/mb.rs nb_motion / pred_mv) -- */

static int nb_motion(sdec_t *t, int l, int x, int y, int *ref, int16_t *mv) {
  int nx, ny;
  const dmb_t *n = nb4(t, x, y, &nx, &ny, 1 + l);
  if (!n) return 0;
  if (n->kind != K_INTER) {
    *ref = -1;
    mv[0] = mv[1] = 0;
  } else {
    *ref = (l ? n->ref1 : n->ref)[(ny >> 1) * 2 + (nx >> 1)];
    mv[0] = (l ? n->mv1 : n->mv)[ny * 4 + nx][0];
    mv[1] = (l ? n->mv1 : n->mv)[ny * 4 + nx][1];
  }
  return 1;
}

static int16_t median3(int16_t a, int16_t b, int16_t c) {
  const int16_t mx = a > b ? a : b, mn = a < b ? a : b;
  const int16_t u = mn > c ? mn : c;
  return mx < u ? mx : u;
}

static void pred_mv(sdec_t *t, int l, int bx, int by, int w, int dir, int rf, int16_t *out) {
  int ra = -1, rb = -1, rc = -1;
  int16_t ma[2] = {0, 0}, mb[2] = {0, 0}, mc[2] = {0, 0};
  const int ha = nb_motion(t, l, bx - 1, by, &ra, ma);
  const int hb = nb_motion(t, l, bx, by - 1, &rb, mb);
  int hc = nb_motion(t, l, bx + w, by - 1, &rc, mc);
  if (!hc) hc = nb_motion(t, l, bx - 1, by - 1, &rc, mc);
  if (!ha) { ra = -1; ma[0] = ma[1] = 0; }
  if (!hb) { rb = -1; mb[0] = mb[1] = 0; }
  if (!hc) { rc = -1; mc[0] = mc[1] = 0; }
  if (dir == 1) {
    if (by == 0 && rb == rf) { out[0] = mb[0]; out[1] = mb[1]; return; }
    if (by != 0 && ra == rf) { out[0] = ma[0]; out[1] = ma[1]; return; }
  } else if (dir == 2) {
    if (bx == 0 && ra == rf) { out[0] = ma[0]; out[1] = ma[1]; return; }
    if (bx != 0 && rc == rf) { out[0] = mc[0]; out[1] = mc[1]; return; }
  }
  if (!hb && !hc && ha) {
    rb = rc = ra;
    mb[0] = mc[0] = ma[0];
    mb[1] = mc[1] = ma[1];
  }
  const int ea = ra == rf, eb = rb == rf, ec = rc == rf;
  if (ea + eb + ec == 1) {
    const int16_t *m = ea ? ma : eb ? mb : mc;
    out[0] = m[0];
    out[1] = m[1];
    return;
  }
  out[0] = median3(ma[0], mb[0], mc[0]);
  out[1] = median3(ma[1], mb[1], mc[1]);
}

static void set_motion(sdec_t *t, int bx, int by, int w, int h, const int16_t *mv, const int *mvd) {
  for (int y = by; y < by + h; y++)
    for (int x = bx; x < bx + w; x++) {
      t->cur->mv[y * 4 + x][0] = mv[0];
      t->cur->mv[y * 4 + x][1] = mv[1];
      t->cur->mvd[y * 4 + x][0] = (uint8_t)mini(absi(mvd[0]), 127);
      t->cur->mvd[y * 4 + x][1] = (uint8_t)mini(absi(mvd[1]), 127);
      t->done |= 1u << (y * 4 + x);
    }
}

/* -- reconstruction -- */

static uint16_t *ypix(sdec_t *t, int x, int y) { return t->d->py + (size_t)y * t->mb_w * 16 + x; }
static uint16_t *cpix(sdec_t *t, int c, int x, int y) { return (c ? t->d->pcr : t->d->pcb) + (size_t)y * t->mb_w * 8 + x; }

/* availability of a neighbour macroblock for intra prediction (constrained_intra_pred keeps to intra ones) */
static int intra_ok(sdec_t *t, const dmb_t *n) { return n && (!t->q->constrained_intra || n->kind != K_INTER); }

/* LevelScale4x4 (8.5.9) of list `li` */
static int ls4(sdec_t *t, int li, int m, int pos) {
  /* the table's columns: both coordinates even, one odd, both odd */
  return t->q->sl4[li][pos] * OD_DEQUANT4_COEFF_INIT[m][(pos & 1) + ((pos >> 2) & 1)];
}

static int ls8(sdec_t *t, int li, int m, int pos) {
  const int i = pos >> 3, j = pos & 7;
  return t->q->sl8[li][pos] * OD_DEQUANT8_COEFF_INIT[m][OD_DEQUANT8_COEFF_INIT_SCAN[((i & 3) << 2) + (j & 3)]];
}

static void dequant4(sdec_t *t, int32_t *c, int li, int qp, int skip_dc) {
  const int m = qp % 6, e = qp / 6;
  for (int p = skip_dc; p < 16; p++) {
    if (!c[p]) continue;
    const int v = c[p] * ls4(t, li, m, p);
    c[p] = qp >= 24 ? shl(v, e - 4) : (v + (1 << (3 - e))) >> (4 - e);
  }
}

static void dequant8(sdec_t *t, int32_t *c, int li, int qp) {
  const int m = qp % 6, e = qp / 6;
  for (int p = 0; p < 64; p++) {
    if (!c[p]) continue;
    const int v = c[p] * ls8(t, li, m, p);
    c[p] = qp >= 36 ? shl(v, e - 6) : (v + (1 << (5 - e))) >> (6 - e);
  }
}

static int chroma_qp_of(int qp, int off) {
  static const uint8_t T[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
  const int q = clampi(qp + off, 0, 51);
  return q < 30 ? q : T[q - 30];
}

/* the residual of the current macroblock added to the prediction already in the picture */
static void add_luma_residual(sdec_t *t, int intra16) {
  const int x0 = t->mbx * 16, y0 = t->mby * 16, stride = t->mb_w * 16;
  const int li = t->cur->kind == K_INTER ? 3 : 0;
  if (t->cur->t8) {
    for (int b8 = 0; b8 < 4; b8++) {
      if (!((t->cur->cbp >> b8) & 1)) continue;
      dequant8(t, t->coef8[b8], t->cur->kind == K_INTER ? 1 : 0, t->qp);
      or_idct8_add(t->coef8[b8], ypix(t, x0 + (b8 & 1) * 8, y0 + (b8 >> 1) * 8), stride, 255);
    }
    return;
  }
  for (int b = 0; b < 16; b++) {
    const int bx = b & 3, by = b >> 2;
    int any = 0;
    if (!intra16) dequant4(t, t->coef[b], li, t->qp, 0);
    else dequant4(t, t->coef[b], li, t->qp, 1);
    for (int k = 0; k < 16; k++) any |= t->coef[b][k] != 0;
    if (any) or_idct4_add(t->coef[b], ypix(t, x0 + bx * 4, y0 + by * 4), stride, 255);
  }
}

static void intra16_dc(sdec_t *t) {
  /* inverse Hadamard of the DC levels, then their scaling (8.5.10) */
  int32_t f[16];
  or_hadamard4(t->dc, f);
  const int qp = t->qp, m = qp % 6, e = qp / 6, lsd = ls4(t, 0, m, 0);
  for (int b = 0; b < 16; b++) t->coef[b][0] = qp >= 36 ? shl(f[b] * lsd, e - 6) : (f[b] * lsd + (1 << (5 - e))) >> (6 - e);
}

static void add_chroma_residual(sdec_t *t) {
  const int inter = t->cur->kind == K_INTER;
  for (int c = 0; c < 2; c++) {
    const int qpc = chroma_qp_of(t->qp, t->q->chroma_qp_offset[c]), m = qpc % 6, e = qpc / 6;
    const int li = (inter ? 3 : 0) + 1 + c;
    const int32_t *l = t->cdc[c];
    const int f[4] = {l[0] + l[1] + l[2] + l[3], l[0] - l[1] + l[2] - l[3], l[0] + l[1] - l[2] - l[3], l[0] - l[1] - l[2] + l[3]};
    for (int blk = 0; blk < 4; blk++) {
      int32_t co[16];
      for (int k = 0; k < 16; k++) co[k] = t->cac[c][blk][k];
      dequant4(t, co, li, qpc, 1);
      co[0] = shl(f[blk] * ls4(t, li, m, 0), e) >> 5;
      int any = 0;
      for (int k = 0; k < 16; k++) any |= co[k] != 0;
      if (any) or_idct4_add(co, cpix(t, c, t->mbx * 8 + (blk & 1) * 4, t->mby * 8 + (blk >> 1) * 4), t->mb_w * 8, 255);
    }
  }
}

static void intra_chroma_pred(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t), *dd = mb_at(t, t->mbx - 1, t->mby - 1);
  const int ha = intra_ok(t, a), hb = intra_ok(t, b), hd = intra_ok(t, dd);
  for (int c = 0; c < 2; c++) {
    int32_t top[8] = {0}, left[16] = {0}, tl = 0;
    const int x0 = t->mbx * 8, y0 = t->mby * 8;
    if (hb)
      for (int i = 0; i < 8; i++) top[i] = *cpix(t, c, x0 + i, y0 - 1);
    if (ha)
      for (int j = 0; j < 8; j++) left[j] = *cpix(t, c, x0 - 1, y0 + j);
    if (hd) tl = *cpix(t, c, x0 - 1, y0 - 1);
    uint16_t out[64];
    or_predict_chroma(t->cur->chroma_mode, top, left, tl, hb, ha, out, 8, 8);
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) *cpix(t, c, x0 + x, y0 + y) = out[y * 8 + x];
  }
}

static int zorder(int bx, int by) { return ((by >> 1) << 3) | ((bx >> 1) << 2) | ((by & 1) << 1) | (bx & 1); }

static void block_pos(int n, int *bx, int *by) {
  *bx = ((n >> 2) & 1) * 2 + (n & 1);
  *by = ((n >> 3) & 1) * 2 + ((n >> 1) & 1);
}

/* Intra4x4PredMode / Intra8x8PredMode derivation (8.3.1.1) */
static int pred_intra_mode(sdec_t *t, int bx, int by) {
  int a = -1, b = -1, nx, ny;
  const dmb_t *na = nb4(t, bx - 1, by, &nx, &ny, 0);
  if (na && (!t->q->constrained_intra || na->kind != K_INTER))
    a = (na->kind == K_I4 || na->kind == K_I8) ? na->modes[ny * 4 + nx] : 2;
  else if (na) a = -1;
  const dmb_t *nbb = nb4(t, bx, by - 1, &nx, &ny, 0);
  if (nbb && (!t->q->constrained_intra || nbb->kind != K_INTER))
    b = (nbb->kind == K_I4 || nbb->kind == K_I8) ? nbb->modes[ny * 4 + nx] : 2;
  if (a < 0 || b < 0) return 2;
  return a < b ? a : b;
}

static void intra_nxn(sdec_t *t, int n8, int bx, int by, int mode) {
  /* edges of the block at 4x4 position (bx, by), n = 4 or 8 samples */
  const int n = n8 ? 8 : 4, x = t->mbx * 16 + bx * 4, y = t->mby * 16 + by * 4;
  const dmb_t *a = left_mb(t), *b = top_mb(t), *dd = mb_at(t, t->mbx - 1, t->mby - 1),
              *cc = mb_at(t, t->mbx + 1, t->mby - 1);
  const int la = intra_ok(t, a), ta = intra_ok(t, b), tla = intra_ok(t, dd), tra = intra_ok(t, cc);
  const int have_left = bx > 0 || la, have_top = by > 0 || ta;
  int have_tl;
  if (bx > 0 && by > 0) have_tl = 1;
  else if (by > 0) have_tl = la;
  else if (bx > 0) have_tl = ta;
  else have_tl = tla;
  int have_tr;
  if (n8) {
    const int i8 = (by >> 1) * 2 + (bx >> 1);
    have_tr = i8 == 0 ? ta : i8 == 1 ? tra : i8 == 2 ? 1 : 0;
  } else if (by == 0) {
    have_tr = bx < 3 ? ta : tra;
  } else if (bx == 3) {
    have_tr = 0;
  } else {
    have_tr = zorder(bx + 1, by - 1) < zorder(bx, by);
  }
  int32_t top[16] = {0}, left[8] = {0}, tl = 0;
  if (have_top) {
    for (int i = 0; i < n; i++) top[i] = *ypix(t, x + i, y - 1);
    for (int i = n; i < 2 * n; i++) top[i] = have_tr ? *ypix(t, x + i, y - 1) : top[n - 1];
  }
  if (have_left)
    for (int j = 0; j < n; j++) left[j] = *ypix(t, x - 1, y + j);
  if (have_tl) tl = *ypix(t, x - 1, y - 1);
  if (n8) or_filter_8x8_edges(top, left, &tl, have_top, have_left, have_tl);
  uint16_t out[64];
  or_predict_nxn(n, mode, top, left, tl, have_top, have_left, have_tl, out, 8);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) *ypix(t, x + c, y + r) = out[r * n + c];
}

/* inter prediction of one partition from reference `ri` (8.4.2), with explicit weights when the PPS
 * asks for them (8.4.2.3) */
static void mc_part(sdec_t *t, int bx, int by, int w, int h, int ri, const int16_t *mv) {
  const frame_t *f = &t->d->frames[t->refs[ri]];
  const int x = t->mbx * 16 + bx * 4, y = t->mby * 16 + by * 4, W = w * 4, H = h * 4;
  const int pw = t->mb_w * 16, ph = t->mb_h * 16;
  uint16_t out[256];
  or_luma_mc_ref(f->y16, pw, pw, ph, x, y, mv[0], mv[1], W, H, out, 255);
  const int wp = t->q->weighted_pred;
  for (int r = 0; r < H; r++)
    for (int c = 0; c < W; c++) {
      int v = out[r * W + c];
      if (wp) {
        const int lwd = t->h->luma_log2wd;
        v = lwd >= 1 ? ((v * t->h->lw[ri] + (1 << (lwd - 1))) >> lwd) + t->h->lo[ri] : v * t->h->lw[ri] + t->h->lo[ri];
        v = clampi(v, 0, 255);
      }
      *ypix(t, x + c, y + r) = (uint16_t)v;
    }
  const int cx = t->mbx * 8 + bx * 2, cy = t->mby * 8 + by * 2, cw = w * 2, chh = h * 2;
  for (int c = 0; c < 2; c++) {
    uint16_t co[64];
    or_chroma_mc(c ? f->cr16 : f->cb16, pw / 2, ph / 2, cx, cy, mv[0], mv[1], cw, chh, co);
    for (int r = 0; r < chh; r++)
      for (int k = 0; k < cw; k++) {
        int v = co[r * cw + k];
        if (wp) {
          const int cwd = t->h->chroma_log2wd;
          v = cwd >= 1 ? ((v * t->h->cw[ri][c] + (1 << (cwd - 1))) >> cwd) + t->h->co[ri][c]
                       : v * t->h->cw[ri][c] + t->h->co[ri][c];
          v = clampi(v, 0, 255);
        }
        *cpix(t, c, cx + k, cy + r) = (uint16_t)v;
      }
  }
}

/* -- B slices (h264/mb.rs decode_inter_mb, direct_spatial/direct_temporal, mc_block; 7.3.5, 8.4.1.2,
 *    8.4.2.3) -- */

/* B mb_type (9.3.2.5, Table 9-37, contexts 27-35): 0..22 the B types, 23 + I mb_type for intra */
static int read_b_mb_type(sdec_t *t) {
  const dmb_t *a = left_mb(t), *b = top_mb(t);
  const int ca = a && !a->bdirect, cb = b && !b->bdirect;
  if (!decision(&t->c, 27 + ca + cb)) return 0;
  if (!decision(&t->c, 30)) return 1 + decision(&t->c, 32);
  int bits = decision(&t->c, 31) << 3;
  bits |= decision(&t->c, 32) << 2;
  bits |= decision(&t->c, 32) << 1;
  bits |= decision(&t->c, 32);
  if (bits <= 7) return bits + 3;
  if (bits == 13) return 23 + read_i_mb_type_at(t, 0, 32);
  if (bits == 14) return 11;
  if (bits == 15) return 22;
  return ((bits << 1) | decision(&t->c, 32)) - 4;
}

/* B sub_mb_type (Table 9-38, contexts 36-39): 0..12 */
static int read_b_sub_type(sdec_t *t) {
  if (!decision(&t->c, 36)) return 0;
  if (!decision(&t->c, 37)) return 1 + decision(&t->c, 39);
  int v = 3;
  if (decision(&t->c, 38)) {
    if (decision(&t->c, 39)) return 11 + decision(&t->c, 39);
    v += 4;
  }
  v += 2 * decision(&t->c, 39);
  v += decision(&t->c, 39);
  return v;
}

/* ref_idx_lX (9.3.3.1.1.6): a neighbouring partition counts when it is inter, not direct, and its
 * index in the list is above 0 */
static int read_ref_idx_l(sdec_t *t, int l, int bx, int by, int max) {
  int cond[2];
  for (int k = 0; k < 2; k++) {
    int nx, ny;
    const dmb_t *n = nb4(t, bx - (k == 0), by - (k == 1), &nx, &ny, 0);
    const int b8 = (ny >> 1) * 2 + (nx >> 1);
    cond[k] = n && n->kind == K_INTER && !((n->direct8 >> b8) & 1) && (l ? n->ref1 : n->ref)[b8] > 0;
  }
  int r = 0, ctx = 54 + cond[0] + 2 * cond[1];
  while (decision(&t->c, ctx)) {
    r++;
    ctx = r == 1 ? 58 : 59;
    if (r > max) return -1;
  }
  return r;
}

static int read_mvd_l(sdec_t *t, int l, int bx, int by, int comp) {
  int nx, ny, sum = 0;
  const dmb_t *a = nb4(t, bx - 1, by, &nx, &ny, 0);
  if (a) sum += (l ? a->mvd1 : a->mvd)[ny * 4 + nx][comp];
  const dmb_t *b = nb4(t, bx, by - 1, &nx, &ny, 0);
  if (b) sum += (l ? b->mvd1 : b->mvd)[ny * 4 + nx][comp];
  const int base = comp ? 47 : 40;
  const int inc = sum < 3 ? 0 : (sum > 32 ? 2 : 1);
  if (!decision(&t->c, base + inc)) return 0;
  int v = 1, ctx = base + 3;
  while (v < 9 && decision(&t->c, ctx)) {
    v++;
    if (ctx < base + 6) ctx++;
  }
  if (v >= 9) v += (int)exp_golomb_bypass(&t->c, 3);
  return bypass(&t->c) ? -v : v;
}

static void set_motion_l(sdec_t *t, int l, int bx, int by, int w, int h, const int16_t *mv) {
  for (int y = by; y < by + h; y++)
    for (int x = bx; x < bx + w; x++) {
      int16_t *d = (l ? t->cur->mv1 : t->cur->mv)[y * 4 + x];
      d[0] = mv[0];
      d[1] = mv[1];
      if (l) t->done1 |= 1u << (y * 4 + x);
      else t->done |= 1u << (y * 4 + x);
    }
}

/* The motion of the co-located 4x4 block (bx, by) in list 1's first picture (h264/mb.rs col_block):
 * its first used list's index, vector and picture; 0 when it is intra. */
static int col_block(sdec_t *t, int bx, int by, int *ref, int16_t *mv, int *id) {
  const frame_t *f = &t->d->frames[t->refs1[0]];
  const ia_mbinfo *m = &f->mbs[t->addr];
  const int cx = t->s->direct8x8 ? (bx >> 1) * 3 : bx, cy = t->s->direct8x8 ? (by >> 1) * 3 : by;
  if (m->kind != K_INTER) return 0;
  const int b8 = (cy >> 1) * 2 + (cx >> 1);
  if (m->ref_idx[b8] >= 0) {
    *ref = m->ref_idx[b8];
    mv[0] = m->mv[cy * 4 + cx][0];
    mv[1] = m->mv[cy * 4 + cx][1];
    *id = m->ref_id[b8];
    return 1;
  }
  if (m->ref_idx1[b8] >= 0) {
    *ref = m->ref_idx1[b8];
    mv[0] = m->mv1[cy * 4 + cx][0];
    mv[1] = m->mv1[cy * 4 + cx][1];
    *id = m->ref_id1[b8];
    return 1;
  }
  return 0;
}

typedef struct {
  int ref[2][4];
  int16_t mv[2][16][2];
} direct_t;

static int min_positive(int x, int y) { return x >= 0 && y >= 0 ? mini(x, y) : (x > y ? x : y); }

/* Spatial direct prediction (8.4.1.2.2) */
static int direct_spatial(sdec_t *t, direct_t *d) {
  int refs[2];
  for (int l = 0; l < 2; l++) {
    int ra = -1, rb = -1, rc = -1, r;
    int16_t m[2];
    if (nb_motion(t, l, -1, 0, &r, m)) ra = r;
    if (nb_motion(t, l, 0, -1, &r, m)) rb = r;
    if (nb_motion(t, l, 4, -1, &r, m)) rc = r;
    else if (nb_motion(t, l, -1, -1, &r, m)) rc = r;
    refs[l] = min_positive(ra, min_positive(rb, rc));
  }
  memset(d, 0, sizeof *d);
  for (int l = 0; l < 2; l++)
    for (int k = 0; k < 4; k++) d->ref[l][k] = -1;
  if (refs[0] < 0 && refs[1] < 0) {
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 4; k++) d->ref[l][k] = 0;
    return OD_OK;
  }
  for (int l = 0; l < 2; l++)
    if (refs[l] >= 0) {
      int16_t mv[2];
      pred_mv(t, l, 0, 0, 4, 0, refs[l], mv);
      for (int k = 0; k < 4; k++) d->ref[l][k] = refs[l];
      for (int b = 0; b < 16; b++) { d->mv[l][b][0] = mv[0]; d->mv[l][b][1] = mv[1]; }
    }
  if (!t->d->frames[t->refs1[0]].long_term)
    for (int b = 0; b < 16; b++) {
      int r, id;
      int16_t mv[2];
      if (col_block(t, b % 4, b / 4, &r, mv, &id) && r == 0 && absi(mv[0]) <= 1 && absi(mv[1]) <= 1)
        for (int l = 0; l < 2; l++)
          if (refs[l] == 0) d->mv[l][b][0] = d->mv[l][b][1] = 0;
    }
  return OD_OK;
}

/* Temporal direct prediction (8.4.1.2.3) */
static int direct_temporal(sdec_t *t, direct_t *d) {
  const frame_t *e1 = &t->d->frames[t->refs1[0]];
  memset(d, 0, sizeof *d);
  for (int i8 = 0; i8 < 4; i8++) {
    const int bx0 = (i8 & 1) * 2, by0 = (i8 >> 1) * 2;
    int r, id, ref0 = 0;
    int16_t mv[2];
    if (col_block(t, bx0, by0, &r, mv, &id)) {
      ref0 = -1;
      for (int i = 0; i < t->h->num_ref_idx_active && ref0 < 0; i++)
        if (t->d->frames[t->refs[i]].id == id) ref0 = i;
      if (ref0 < 0) return fail(t->d, OD_EBITSTREAM, "temporal direct names a picture not in list 0");
    }
    const frame_t *e0 = &t->d->frames[t->refs[ref0]];
    d->ref[0][i8] = ref0;
    d->ref[1][i8] = 0;
    const int tb = clampi(t->d->cur_poc - e0->poc, -128, 127), td = clampi(e1->poc - e0->poc, -128, 127);
    for (int j = 0; j < 4; j++) {
      const int bx = bx0 + (j & 1), by = by0 + (j >> 1);
      int16_t col[2] = {0, 0};
      if (!col_block(t, bx, by, &r, col, &id)) col[0] = col[1] = 0;
      int16_t m0[2], m1[2];
      if (e0->long_term || td == 0) {
        m0[0] = col[0];
        m0[1] = col[1];
        m1[0] = m1[1] = 0;
      } else {
        const int tx = (16384 + absi(td / 2)) / td;
        const int dsf = clampi((tb * tx + 32) >> 6, -1024, 1023);
        for (int c = 0; c < 2; c++) {
          m0[c] = (int16_t)((dsf * col[c] + 128) >> 8);
          m1[c] = (int16_t)(m0[c] - col[c]);
        }
      }
      const int b = by * 4 + bx;
      d->mv[0][b][0] = m0[0];
      d->mv[0][b][1] = m0[1];
      d->mv[1][b][0] = m1[0];
      d->mv[1][b][1] = m1[1];
    }
  }
  return OD_OK;
}

/* The direct motion of the 8x8 blocks in `mask` into the macroblock (h264/mb.rs put_direct) */
static int put_direct(sdec_t *t, int mask) {
  direct_t d;
  const int rc = t->h->direct_spatial ? direct_spatial(t, &d) : direct_temporal(t, &d);
  if (rc) return rc;
  if (!t->h->direct_spatial) t->d->stats[OD_STAT_B_TEMPORAL]++;
  dmb_t *m = t->cur;
  for (int i8 = 0; i8 < 4; i8++) {
    if (!((mask >> i8) & 1)) continue;
    m->ref[i8] = (int8_t)d.ref[0][i8];
    m->ref1[i8] = (int8_t)d.ref[1][i8];
    for (int j = 0; j < 4; j++) {
      const int b = ((i8 >> 1) * 2 + (j >> 1)) * 4 + (i8 & 1) * 2 + (j & 1);
      for (int c = 0; c < 2; c++) {
        m->mv[b][c] = d.ref[0][i8] >= 0 ? d.mv[0][b][c] : 0;
        m->mv1[b][c] = d.ref[1][i8] >= 0 ? d.mv[1][b][c] : 0;
      }
    }
  }
  return OD_OK;
}

static int clip255(int v) { return v < 0 ? 0 : v > 255 ? 255 : v; }

/* The prediction of 4x4 block (bx, by) of a B macroblock from the lists its 8x8 block uses: averaged,
 * or weighted implicitly (weighted_bipred_idc 2) or explicitly (idc 1) (8.4.2.3; h264/mb.rs mc_block) */
static int mc_block_b(sdec_t *t, int bx, int by) {
  const dmb_t *m = t->cur;
  const int b8 = (by >> 1) * 2 + (bx >> 1), b = by * 4 + bx;
  const int x = t->mbx * 16 + bx * 4, y = t->mby * 16 + by * 4, pw = t->mb_w * 16, ph = t->mb_h * 16;
  const int cx = t->mbx * 8 + bx * 2, cy = t->mby * 8 + by * 2;
  uint16_t lp[2][16], cp[2][2][4];
  int have[2] = {0, 0}, poc[2] = {0, 0}, lt[2] = {0, 0}, ri[2] = {0, 0};
  for (int l = 0; l < 2; l++) {
    const int r = (l ? m->ref1 : m->ref)[b8];
    if (r < 0) continue;
    const int na = l ? t->h->num_ref_idx_active1 : t->h->num_ref_idx_active;
    if (r >= na) return fail(t->d, OD_EBITSTREAM, "reference index past its list");
    const frame_t *f = &t->d->frames[(l ? t->refs1 : t->refs)[r]];
    const int16_t *mv = (l ? m->mv1 : m->mv)[b];
    or_luma_mc_ref(f->y16, pw, pw, ph, x, y, mv[0], mv[1], 4, 4, lp[l], 255);
    for (int c = 0; c < 2; c++) or_chroma_mc(c ? f->cr16 : f->cb16, pw / 2, ph / 2, cx, cy, mv[0], mv[1], 2, 2, cp[l][c]);
    have[l] = 1;
    poc[l] = f->poc;
    lt[l] = f->long_term;
    ri[l] = r;
  }
  if (!have[0] && !have[1]) return fail(t->d, OD_EBITSTREAM, "inter block without a reference");
  const sh_t *h = t->h;
  const int explicit_w = t->q->weighted_bipred == 1;
  uint16_t outl[16], outc[2][4];
  if (have[0] && have[1]) {
    /* (log2 denominator, w0, w1, offset) for luma and each chroma plane */
    int wt[3][4] = {{0}};
    int weighted = 0;
    if (explicit_w) {
      weighted = 1;
      wt[0][0] = h->luma_log2wd;
      wt[0][1] = h->lw[ri[0]];
      wt[0][2] = h->lw1[ri[1]];
      wt[0][3] = (h->lo[ri[0]] + h->lo1[ri[1]] + 1) >> 1;
      for (int c = 0; c < 2; c++) {
        wt[1 + c][0] = h->chroma_log2wd;
        wt[1 + c][1] = h->cw[ri[0]][c];
        wt[1 + c][2] = h->cw1[ri[1]][c];
        wt[1 + c][3] = (h->co[ri[0]][c] + h->co1[ri[1]][c] + 1) >> 1;
      }
      t->d->stats[OD_STAT_B_EXPLICIT]++;
    } else if (t->q->weighted_bipred == 2) {
      weighted = 1;
      int w0 = 32, w1 = 32;
      const int tb = clampi(t->d->cur_poc - poc[0], -128, 127), td = clampi(poc[1] - poc[0], -128, 127);
      if (td != 0 && !lt[0] && !lt[1]) {
        const int tx = (16384 + absi(td / 2)) / td;
        const int dsf = clampi((tb * tx + 32) >> 6, -1024, 1023);
        if ((dsf >> 2) >= -64 && (dsf >> 2) <= 128) {
          w1 = dsf >> 2;
          w0 = 64 - w1;
        }
      }
      for (int k = 0; k < 3; k++) {
        wt[k][0] = 5;
        wt[k][1] = w0;
        wt[k][2] = w1;
        wt[k][3] = 0;
      }
      t->d->stats[OD_STAT_B_IMPLICIT]++;
    }
    for (int i = 0; i < 16; i++)
      outl[i] = weighted ? (uint16_t)clip255(((lp[0][i] * wt[0][1] + lp[1][i] * wt[0][2] + (1 << wt[0][0])) >> (wt[0][0] + 1)) + wt[0][3])
                         : (uint16_t)((lp[0][i] + lp[1][i] + 1) >> 1);
    for (int c = 0; c < 2; c++)
      for (int i = 0; i < 4; i++)
        outc[c][i] = weighted ? (uint16_t)clip255(((cp[0][c][i] * wt[1 + c][1] + cp[1][c][i] * wt[1 + c][2] + (1 << wt[1 + c][0])) >> (wt[1 + c][0] + 1)) + wt[1 + c][3])
                              : (uint16_t)((cp[0][c][i] + cp[1][c][i] + 1) >> 1);
  } else {
    const int l = have[0] ? 0 : 1, r = ri[l];
    for (int i = 0; i < 16; i++) {
      int v = lp[l][i];
      if (explicit_w) {
        const int lwd = h->luma_log2wd, w = (l ? h->lw1 : h->lw)[r], o = (l ? h->lo1 : h->lo)[r];
        v = clip255(lwd >= 1 ? ((v * w + (1 << (lwd - 1))) >> lwd) + o : v * w + o);
      }
      outl[i] = (uint16_t)v;
    }
    for (int c = 0; c < 2; c++)
      for (int i = 0; i < 4; i++) {
        int v = cp[l][c][i];
        if (explicit_w) {
          const int cwd = h->chroma_log2wd, w = (l ? h->cw1 : h->cw)[r][c], o = (l ? h->co1 : h->co)[r][c];
          v = clip255(cwd >= 1 ? ((v * w + (1 << (cwd - 1))) >> cwd) + o : v * w + o);
        }
        outc[c][i] = (uint16_t)v;
      }
  }
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) *ypix(t, x + c, y + r) = outl[r * 4 + c];
  for (int c = 0; c < 2; c++)
    for (int r = 0; r < 2; r++)
      for (int k = 0; k < 2; k++) *cpix(t, c, cx + k, cy + r) = outc[c][r * 2 + k];
  return OD_OK;
}

static int mc_mb_b(sdec_t *t) {
  for (int b = 0; b < 16; b++) {
    const int rc = mc_block_b(t, b % 4, b / 4);
    if (rc) return rc;
  }
  return OD_OK;
}

static void reset_motion(dmb_t *m) {
  for (int i = 0; i < 4; i++) m->ref[i] = m->ref1[i] = -1;
  memset(m->mv, 0, sizeof m->mv);
  memset(m->mv1, 0, sizeof m->mv1);
  memset(m->mvd, 0, sizeof m->mvd);
  memset(m->mvd1, 0, sizeof m->mvd1);
}

/* B_Skip: direct prediction of the whole macroblock, no residual */
static int decode_b_skip(sdec_t *t) {
  dmb_t *m = t->cur;
  m->kind = K_INTER;
  m->skip = 1;
  m->cbp = 0;
  reset_motion(m);
  m->bdirect = 1;
  m->direct8 = 15;
  t->d->stats[OD_STAT_B_SKIP]++;
  int rc = put_direct(t, 15);
  if (!rc) rc = mc_mb_b(t);
  t->prev_qp_delta_nz = 0;
  return rc;
}

static int read_residual(sdec_t *t, int intra);
static void add_luma_residual(sdec_t *t, int intra16);
static void add_chroma_residual(sdec_t *t);

/* An inter macroblock of a B slice, mb_type bt in 0..22 (Table 7-14) */
static int decode_b_inter(sdec_t *t, int bt) {
  static const int SUBM[13][2] = {{0, 0}, {1, 0}, {2, 0}, {3, 0}, {1, 1}, {1, 2}, {2, 1}, {2, 2}, {3, 1}, {3, 2},
                                  {1, 3}, {2, 3}, {3, 3}}; /* (list mask, shape 0 8x8 1 8x4 2 4x8 3 4x4) */
  static const uint8_t PAIRS[9][2] = {{1, 1}, {2, 2}, {1, 2}, {2, 1}, {1, 3}, {2, 3}, {3, 1}, {3, 2}, {3, 3}};
  dmb_t *m = t->cur;
  m->kind = K_INTER;
  m->chroma_mode = 0;
  reset_motion(m);
  int rp[4][5], nrp = 0;            /* (bx, by, w, h, mask) */
  int mp[16][7], nmp = 0;           /* (bx, by, w, h, mask, dir, direct) */
  int direct_mask = 0, sub8_ok = 1;
#define RP(x, y, w, h, k) do { rp[nrp][0] = x; rp[nrp][1] = y; rp[nrp][2] = w; rp[nrp][3] = h; rp[nrp][4] = k; nrp++; } while (0)
#define MP(x, y, w, h, k, d, dr) do { mp[nmp][0] = x; mp[nmp][1] = y; mp[nmp][2] = w; mp[nmp][3] = h; mp[nmp][4] = k; mp[nmp][5] = d; mp[nmp][6] = dr; nmp++; } while (0)
  if (bt == 0) {
    direct_mask = 15;
    m->bdirect = 1;
    t->d->stats[OD_STAT_B_DIRECT16]++;
  } else if (bt <= 3) {
    RP(0, 0, 4, 4, bt);
    MP(0, 0, 4, 4, bt, 0, 0);
    t->d->stats[OD_STAT_B16 + bt - 1]++;
  } else if (bt < 22) {
    const int idx = bt - 4, m0 = PAIRS[idx / 2][0], m1 = PAIRS[idx / 2][1];
    if (idx % 2 == 0) {
      RP(0, 0, 4, 2, m0); RP(0, 2, 4, 2, m1);
      MP(0, 0, 4, 2, m0, 1, 0); MP(0, 2, 4, 2, m1, 1, 0);
    } else {
      RP(0, 0, 2, 4, m0); RP(2, 0, 2, 4, m1);
      MP(0, 0, 2, 4, m0, 2, 0); MP(2, 0, 2, 4, m1, 2, 0);
    }
    t->d->stats[OD_STAT_B_PART + idx % 2]++;
  } else {
    int sub[4];
    for (int k = 0; k < 4; k++) {
      sub[k] = read_b_sub_type(t);
      t->d->stats[OD_STAT_B_SUB + sub[k]]++;
    }
    t->d->stats[OD_STAT_B8X8]++;
    for (int i = 0; i < 4; i++) {
      const int x0 = (i & 1) * 2, y0 = (i >> 1) * 2;
      if (sub[i] == 0) {
        direct_mask |= 1 << i;
        MP(x0, y0, 2, 2, 0, 0, 1);
        continue;
      }
      const int mask = SUBM[sub[i]][0], shape = SUBM[sub[i]][1];
      sub8_ok &= shape == 0;
      RP(x0, y0, 2, 2, mask);
      static const int sp[4][4][4] = {{{0, 0, 2, 2}}, {{0, 0, 2, 1}, {0, 1, 2, 1}}, {{0, 0, 1, 2}, {1, 0, 1, 2}},
                                      {{0, 0, 1, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}}};
      static const int nsp[4] = {1, 2, 2, 4};
      for (int q = 0; q < nsp[shape]; q++) MP(x0 + sp[shape][q][0], y0 + sp[shape][q][1], sp[shape][q][2], sp[shape][q][3], mask, 0, 0);
    }
  }
#undef RP
#undef MP
  /* direct motion first: later partitions may predict from it */
  m->direct8 = (uint8_t)direct_mask;
  if (direct_mask) {
    const int rc = put_direct(t, direct_mask);
    if (rc) return rc;
  }
  const int maxr[2] = {t->h->num_ref_idx_active - 1, t->h->num_ref_idx_active1 - 1};
  for (int l = 0; l < 2; l++)
    for (int k = 0; k < nrp; k++) {
      if (!((rp[k][4] >> l) & 1)) continue;
      const int r = maxr[l] > 0 ? read_ref_idx_l(t, l, rp[k][0], rp[k][1], maxr[l]) : 0;
      if (r < 0) return fail(t->d, OD_EBITSTREAM, "ref_idx past its list");
      for (int yy = rp[k][1]; yy < rp[k][1] + rp[k][3]; yy += 2)
        for (int xx = rp[k][0]; xx < rp[k][0] + rp[k][2]; xx += 2) (l ? m->ref1 : m->ref)[(yy >> 1) * 2 + (xx >> 1)] = (int8_t)r;
    }
  for (int l = 0; l < 2; l++) {
    if (l) t->done1 = 0;
    else t->done = 0;
    for (int k = 0; k < nmp; k++) {
      const int bx = mp[k][0], by = mp[k][1], w = mp[k][2], h = mp[k][3];
      if (mp[k][6]) {
        for (int yy = by; yy < by + h; yy++)
          for (int xx = bx; xx < bx + w; xx++) {
            if (l) t->done1 |= 1u << (yy * 4 + xx);
            else t->done |= 1u << (yy * 4 + xx);
          }
      } else if ((mp[k][4] >> l) & 1) {
        const int refi = (l ? m->ref1 : m->ref)[(by >> 1) * 2 + (bx >> 1)];
        int16_t p[2], mv[2];
        pred_mv(t, l, bx, by, w, mp[k][5], refi, p);
        const int d0 = read_mvd_l(t, l, bx, by, 0), d1 = read_mvd_l(t, l, bx, by, 1);
        for (int yy = by; yy < by + h; yy++)
          for (int xx = bx; xx < bx + w; xx++) {
            uint8_t *md = (l ? m->mvd1 : m->mvd)[yy * 4 + xx];
            md[0] = (uint8_t)mini(absi(d0), 70);
            md[1] = (uint8_t)mini(absi(d1), 70);
          }
        mv[0] = (int16_t)(p[0] + d0);
        mv[1] = (int16_t)(p[1] + d1);
        set_motion_l(t, l, bx, by, w, h, mv);
      } else {
        const int16_t z[2] = {0, 0};
        set_motion_l(t, l, bx, by, w, h, z);
      }
    }
  }
  m->cbp = (uint8_t)read_cbp(t);
  const int direct_ok = direct_mask == 0 || t->s->direct8x8;
  if ((m->cbp & 15) && t->q->transform8 && sub8_ok && direct_ok) m->t8 = (uint8_t)read_t8_flag(t);
  if (m->t8) t->d->stats[OD_STAT_INTER_T8]++;
  int qd = 0;
  if (m->cbp) qd = read_qp_delta(t);
  t->prev_qp_delta_nz = qd != 0;
  t->qp = (t->qp + qd + 52) % 52;
  m->qp = t->qp;
  read_residual(t, 0);
  const int rc = mc_mb_b(t);
  if (rc) return rc;
  add_luma_residual(t, 0);
  add_chroma_residual(t);
  return OD_OK;
}

/* -- macroblock layer -- */

static int read_residual(sdec_t *t, int intra) {
  dmb_t *m = t->cur;
  const int cbpl = m->cbp & 15, cbpc = m->cbp >> 4;
  memset(t->coef, 0, sizeof t->coef);
  memset(t->coef8, 0, sizeof t->coef8);
  memset(t->cdc, 0, sizeof t->cdc);
  memset(t->cac, 0, sizeof t->cac);
  memset(m->nz, 0, sizeof m->nz);
  memset(m->cnz, 0, sizeof m->cnz);
  m->cbf_dc = 0;
  if (m->kind == K_I16) {
    int32_t lv[16];
    const int n = read_block(t, 0, cbf_dc(t, 0, 1, 0), 16, lv);
    if (n) m->cbf_dc |= 1;
    for (int k = 0; k < 16; k++) t->dc[OD_ZIGZAG4[k]] = lv[k];
  }
  for (int b8 = 0; b8 < 4; b8++) {
    if (!((cbpl >> b8) & 1)) continue;
    if (m->t8) {
      int32_t lv[64];
      const int n = read_block(t, 5, -1, 64, lv);
      for (int k = 0; k < 64; k++) t->coef8[b8][OD_ZIGZAG_DIRECT[k]] = lv[k];
      const int bx = (b8 & 1) * 2, by = (b8 >> 1) * 2;
      for (int j = 0; j < 4; j++) m->nz[(by + (j >> 1)) * 4 + bx + (j & 1)] = (uint8_t)n;
      continue;
    }
    for (int b4 = 0; b4 < 4; b4++) {
      int bx, by;
      block_pos(b8 * 4 + b4, &bx, &by);
      int32_t lv[16];
      const int i16 = m->kind == K_I16;
      const int inc = cbf_luma4(t, bx, by, intra);
      const int n = read_block(t, i16 ? 1 : 2, inc, i16 ? 15 : 16, lv);
      int32_t *c = t->coef[by * 4 + bx];
      if (i16)
        for (int k = 0; k < 15; k++) c[OD_ZIGZAG4[k + 1]] = lv[k];
      else
        for (int k = 0; k < 16; k++) c[OD_ZIGZAG4[k]] = lv[k];
      m->nz[by * 4 + bx] = (uint8_t)n;
    }
  }
  if (cbpc) {
    for (int c = 0; c < 2; c++) {
      const int n = read_block(t, 3, cbf_dc(t, 1 + c, intra, 1), 4, t->cdc[c]);
      if (n) m->cbf_dc |= (uint8_t)(2 << c);
    }
  }
  if (cbpc == 2) {
    for (int c = 0; c < 2; c++)
      for (int blk = 0; blk < 4; blk++) {
        int32_t lv[16];
        const int n = read_block(t, 4, cbf_chroma_ac(t, c, blk & 1, blk >> 1, intra), 15, lv);
        for (int k = 0; k < 15; k++) t->cac[c][blk][OD_ZIGZAG4[k + 1]] = lv[k];
        m->cnz[c][blk] = (uint8_t)n;
      }
  }
  return OD_OK;
}

static int decode_pcm(sdec_t *t) {
  /* pcm_alignment_zero_bits, 384 samples, then the engine starts again (9.3.1.2) */
  br_t *b = t->c.b;
  /* after a terminate bin of 1 the decoder has read exactly the bits the encoder's flush wrote (its 9
     register bits included), so the alignment bits start here (9.3.1.2, 9.3.4.5) */
  while (b->pos & 7) b->pos++;
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) *ypix(t, t->mbx * 16 + x, t->mby * 16 + y) = (uint16_t)br_u(b, 8);
  for (int c = 0; c < 2; c++)
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) *cpix(t, c, t->mbx * 8 + x, t->mby * 8 + y) = (uint16_t)br_u(b, 8);
  cabac_init_engine(&t->c);
  dmb_t *m = t->cur;
  t->d->stats[OD_STAT_PCM]++;
  m->kind = K_PCM;
  m->cbp = 0x2f;
  memset(m->nz, 16, sizeof m->nz);
  memset(m->cnz, 16, sizeof m->cnz);
  m->cbf_dc = 7;
  m->chroma_mode = 0;
  for (int i = 0; i < 4; i++) m->ref[i] = -1;
  t->prev_qp_delta_nz = 0;
  return OD_OK;
}

static int decode_skip(sdec_t *t) {
  /* P_Skip (8.4.1.1) */
  dmb_t *m = t->cur;
  m->kind = K_INTER;
  m->skip = 1;
  t->d->stats[OD_STAT_SKIP]++;
  m->cbp = 0;
  for (int i = 0; i < 4; i++) m->ref[i] = 0;
  int16_t mv[2] = {0, 0};
  int ra, rb;
  int16_t ma[2], mb[2];
  const int ha = nb_motion(t, 0, -1, 0, &ra, ma), hb = nb_motion(t, 0, 0, -1, &rb, mb);
  if (!(!ha || !hb || (ra == 0 && ma[0] == 0 && ma[1] == 0) || (rb == 0 && mb[0] == 0 && mb[1] == 0)))
    pred_mv(t, 0, 0, 0, 4, 0, 0, mv);
  const int z[2] = {0, 0};
  set_motion(t, 0, 0, 4, 4, mv, z);
  memset(m->mvd, 0, sizeof m->mvd);
  mc_part(t, 0, 0, 4, 4, 0, mv);
  t->prev_qp_delta_nz = 0;
  return OD_OK;
}

static int decode_mb(sdec_t *t, int islice) {
  dmb_t *m = t->cur;
  int mt;
  if (islice) mt = read_i_mb_type(t, 1);
  else if (t->bslice) {
    const int bt = read_b_mb_type(t);
    if (bt < 23) return decode_b_inter(t, bt);
    mt = 5 + bt - 23;
  } else mt = read_p_mb_type(t);
  int intra = islice || mt >= 5;
  if (!islice && mt >= 5) mt -= 5;
  else if (!islice) {
    /* P types */
  }
  memset(m->modes, 2, sizeof m->modes);
  if (intra && mt == 25) return decode_pcm(t);
  m->t8 = 0;
  if (intra) {
    reset_motion(m);
    if (mt == 0) {
      if (t->q->transform8) m->t8 = (uint8_t)read_t8_flag(t);
      m->kind = m->t8 ? K_I8 : K_I4;
      int syn[16][2];
      for (int n = 0; n < (m->t8 ? 4 : 16); n++) {
        syn[n][0] = decision(&t->c, 68);
        syn[n][1] = 0;
        if (!syn[n][0])
          for (int k = 0; k < 3; k++) syn[n][1] |= decision(&t->c, 69) << k;
      }
      m->chroma_mode = (uint8_t)read_chroma_mode(t);
      m->cbp = (uint8_t)read_cbp(t);
      int qd = 0;
      if (m->cbp) {
        if (m->t8 == 0 && 0) (void)0;
        qd = read_qp_delta(t);
      }
      t->prev_qp_delta_nz = qd != 0;
      t->qp = (t->qp + qd + 52) % 52;
      m->qp = t->qp;
      read_residual(t, 1);
      /* predict and reconstruct block by block, deriving each mode as it goes */
      for (int n = 0; n < (m->t8 ? 4 : 16); n++) {
        int bx, by;
        if (m->t8) { bx = (n & 1) * 2; by = (n >> 1) * 2; }
        else block_pos(n, &bx, &by);
        const int pred = pred_intra_mode(t, bx, by);
        const int mode = syn[n][0] ? pred : (syn[n][1] < pred ? syn[n][1] : syn[n][1] + 1);
        if (m->t8) {
          for (int j = 0; j < 4; j++) m->modes[(by + (j >> 1)) * 4 + bx + (j & 1)] = (uint8_t)mode;
          t->d->stats[OD_STAT_I8 + mode]++;
          intra_nxn(t, 1, bx, by, mode);
          const int b8 = n;
          if ((m->cbp >> b8) & 1) {
            dequant8(t, t->coef8[b8], 0, t->qp);
            or_idct8_add(t->coef8[b8], ypix(t, t->mbx * 16 + bx * 4, t->mby * 16 + by * 4), t->mb_w * 16, 255);
          }
        } else {
          m->modes[by * 4 + bx] = (uint8_t)mode;
          t->d->stats[OD_STAT_I4 + mode]++;
          intra_nxn(t, 0, bx, by, mode);
          int32_t *c = t->coef[by * 4 + bx];
          dequant4(t, c, 0, t->qp, 0);
          int any = 0;
          for (int k = 0; k < 16; k++) any |= c[k] != 0;
          if (any) or_idct4_add(c, ypix(t, t->mbx * 16 + bx * 4, t->mby * 16 + by * 4), t->mb_w * 16, 255);
        }
      }
      t->d->stats[OD_STAT_CHROMA + m->chroma_mode]++;
      intra_chroma_pred(t);
      add_chroma_residual(t);
      return OD_OK;
    }
    /* Intra16x16: mb_type 1..24 */
    m->kind = K_I16;
    t->i16_mode = (mt - 1) % 4;
    t->d->stats[OD_STAT_I16 + t->i16_mode]++;
    const int cc = ((mt - 1) / 4) % 3, luma = mt >= 13;
    m->chroma_mode = (uint8_t)read_chroma_mode(t);
    m->cbp = (uint8_t)((luma ? 15 : 0) | cc << 4);
    const int qd = read_qp_delta(t);
    t->prev_qp_delta_nz = qd != 0;
    t->qp = (t->qp + qd + 52) % 52;
    m->qp = t->qp;
    read_residual(t, 1);
    {
      const dmb_t *a = left_mb(t), *b = top_mb(t), *dd = mb_at(t, t->mbx - 1, t->mby - 1);
      const int ha = intra_ok(t, a), hb = intra_ok(t, b), hd = intra_ok(t, dd);
      int32_t top[16] = {0}, left[16] = {0}, tl = 0;
      const int x0 = t->mbx * 16, y0 = t->mby * 16;
      if (hb)
        for (int i = 0; i < 16; i++) top[i] = *ypix(t, x0 + i, y0 - 1);
      if (ha)
        for (int j = 0; j < 16; j++) left[j] = *ypix(t, x0 - 1, y0 + j);
      if (hd) tl = *ypix(t, x0 - 1, y0 - 1);
      uint16_t out[256];
      or_predict_16x16(t->i16_mode, top, left, tl, hb, ha, out, 8);
      for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) *ypix(t, x0 + x, y0 + y) = out[y * 16 + x];
    }
    t->d->stats[OD_STAT_CHROMA + m->chroma_mode]++;
    intra16_dc(t);
    add_luma_residual(t, 1);
    intra_chroma_pred(t);
    add_chroma_residual(t);
    return OD_OK;
  }
  /* inter: P_L0_16x16, 16x8, 8x16, P_8x8 */
  m->kind = K_INTER;
  m->chroma_mode = 0;
  const int part = mt;
  int sub[4] = {0, 0, 0, 0};
  t->d->stats[OD_STAT_PART + part]++;
  if (part == 3)
    for (int k = 0; k < 4; k++) {
      sub[k] = read_sub_mb_type(t);
      t->d->stats[OD_STAT_SUB + sub[k]]++;
    }
  const int na = t->h->num_ref_idx_active;
  int refs[4] = {0, 0, 0, 0};
  static const int np[4] = {1, 2, 2, 4};
  static const int pr[4][4][4] = {{{0, 0, 4, 4}}, {{0, 0, 4, 2}, {0, 2, 4, 2}}, {{0, 0, 2, 4}, {2, 0, 2, 4}},
                                  {{0, 0, 2, 2}, {2, 0, 2, 2}, {0, 2, 2, 2}, {2, 2, 2, 2}}};
  for (int k = 0; k < np[part]; k++) {
    if (na > 1) refs[k] = read_ref_idx(t, pr[part][k][0], pr[part][k][1]);
    if (refs[k] >= na) return fail(t->d, OD_EBITSTREAM, "ref_idx %d of %d", refs[k], na);
    if (refs[k] > 0) t->d->stats[OD_STAT_REF_GT0]++;
    /* the partition's 8x8 blocks carry its reference for the next ref_idx contexts */
    for (int y8 = pr[part][k][1] / 2; y8 <= (pr[part][k][1] + pr[part][k][3] - 1) / 2; y8++)
      for (int x8 = pr[part][k][0] / 2; x8 <= (pr[part][k][0] + pr[part][k][2] - 1) / 2; x8++)
        m->ref[y8 * 2 + x8] = (int8_t)refs[k];
  }
  /* mvds, in partition and sub-partition order; vectors predicted as they are decoded */
  memset(m->mvd, 0, sizeof m->mvd);
  t->done = 0;
  static const int sp[4][4][4] = {{{0, 0, 2, 2}}, {{0, 0, 2, 1}, {0, 1, 2, 1}}, {{0, 0, 1, 2}, {1, 0, 1, 2}},
                                  {{0, 0, 1, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}}};
  static const int nsp[4] = {1, 2, 2, 4};
  const int dir = part == 1 ? 1 : part == 2 ? 2 : 0;
  for (int k = 0; k < np[part]; k++) {
    const int n_sub = part == 3 ? nsp[sub[k]] : 1;
    for (int s = 0; s < n_sub; s++) {
      int bx, by, w, h;
      if (part == 3) {
        bx = pr[3][k][0] + sp[sub[k]][s][0];
        by = pr[3][k][1] + sp[sub[k]][s][1];
        w = sp[sub[k]][s][2];
        h = sp[sub[k]][s][3];
      } else {
        bx = pr[part][k][0];
        by = pr[part][k][1];
        w = pr[part][k][2];
        h = pr[part][k][3];
      }
      int mvd[2];
      mvd[0] = read_mvd(t, bx, by, 0);
      mvd[1] = read_mvd(t, bx, by, 1);
      int16_t mvp[2], mv[2];
      pred_mv(t, 0, bx, by, w, part == 3 ? 0 : dir, refs[k], mvp);
      mv[0] = (int16_t)(mvp[0] + mvd[0]);
      mv[1] = (int16_t)(mvp[1] + mvd[1]);
      set_motion(t, bx, by, w, h, mv, mvd);
    }
  }
  m->cbp = (uint8_t)read_cbp(t);
  if ((m->cbp & 15) && t->q->transform8 && (part != 3 || (sub[0] | sub[1] | sub[2] | sub[3]) == 0))
    m->t8 = (uint8_t)read_t8_flag(t);
  if (m->t8) t->d->stats[OD_STAT_INTER_T8]++;
  int qd = 0;
  if (m->cbp) qd = read_qp_delta(t);
  t->prev_qp_delta_nz = qd != 0;
  t->qp = (t->qp + qd + 52) % 52;
  m->qp = t->qp;
  read_residual(t, 0);
  /* prediction of each partition, then the residual */
  for (int k = 0; k < np[part]; k++) {
    if (part == 3) {
      const int n_sub = nsp[sub[k]];
      for (int s = 0; s < n_sub; s++) {
        const int bx = pr[3][k][0] + sp[sub[k]][s][0], by = pr[3][k][1] + sp[sub[k]][s][1];
        mc_part(t, bx, by, sp[sub[k]][s][2], sp[sub[k]][s][3], refs[k], m->mv[by * 4 + bx]);
      }
    } else {
      const int bx = pr[part][k][0], by = pr[part][k][1];
      mc_part(t, bx, by, pr[part][k][2], pr[part][k][3], refs[k], m->mv[by * 4 + bx]);
    }
  }
  add_luma_residual(t, 0);
  add_chroma_residual(t);
  return OD_OK;
}

/* the record the loop filter and the caller read (ia_mbinfo) */
static void finish_info(sdec_t *t) {
  const dmb_t *m = t->cur;
  ia_mbinfo *info = &t->d->info[t->addr];
  memset(info, 0, sizeof *info);
  info->slice = (uint16_t)t->slice_id;
  info->kind = m->kind;
  info->qp = (int8_t)(m->kind == K_PCM ? 0 : m->qp);
  info->cbp = m->cbp;
  info->skip = m->skip;
  info->t8x8 = m->t8;
  for (int i = 0; i < 16; i++) info->nnz[i] = m->nz[i];
  const int inter = m->kind == K_INTER;
  for (int i = 0; i < 16; i++)
    for (int c = 0; c < 2; c++) {
      info->mv[i][c] = inter && m->ref[i / 8 * 2 + (i % 4) / 2] >= 0 ? m->mv[i][c] : 0;
      info->mv1[i][c] = inter && t->bslice && m->ref1[i / 8 * 2 + (i % 4) / 2] >= 0 ? m->mv1[i][c] : 0;
    }
  for (int b8 = 0; b8 < 4; b8++) {
    const int r0 = inter ? m->ref[b8] : -1, r1 = inter && t->bslice ? m->ref1[b8] : -1;
    info->ref_idx[b8] = (int8_t)r0;
    info->ref_id[b8] = r0 >= 0 ? t->d->frames[t->refs[r0]].id : -1;
    info->ref_idx1[b8] = (int8_t)r1;
    info->ref_id1[b8] = r1 >= 0 ? t->d->frames[t->refs1[r1]].id : -1;
  }
  info->bdirect = m->bdirect;
  info->direct8 = m->direct8;
}

static int decode_slice(od_dec *d, const sh_t *h, br_t *b) {
  sdec_t t;
  memset(&t, 0, sizeof t);
  t.d = d;
  t.q = &d->pps[h->pps_id];
  t.s = &d->sps[t.q->sps_id];
  t.h = h;
  t.mb_w = d->mb_w;
  t.mb_h = d->mb_h;
  if (d->nslices >= 4096) return fail(d, OD_EUNSUPPORTED, "too many slices");
  t.slice_id = ++d->nslices;
  d->sfilter[t.slice_id - 1] = (or_slice_filter){h->disable_deblock, h->alpha_offset, h->beta_offset,
                                                   {t.q->chroma_qp_offset[0], t.q->chroma_qp_offset[1]}};
  const int islice = h->slice_type == 2;
  t.bslice = h->slice_type == 1;
  if (t.bslice) {
    int rc = ref_lists_b(d, t.s, h, t.refs, t.refs1);
    if (rc) return rc;
  } else if (!islice) {
    int rc = ref_list(d, t.s, h, t.refs);
    if (rc) return rc;
  }
  /* slice_data: cabac_alignment_one_bit, then the macroblocks */
  while (b->pos & 7) b->pos++;
  t.c.b = b;
  cabac_init_contexts(&t.c, islice, h->cabac_init_idc, h->qp);
  cabac_init_engine(&t.c);
  t.qp = h->qp;
  const int nmb = d->mb_w * d->mb_h;
  int addr = h->first_mb;
  for (;;) {
    if (addr >= nmb) return fail(d, OD_EBITSTREAM, "macroblock address %d past the picture", addr);
    t.addr = addr;
    t.mbx = addr % d->mb_w;
    t.mby = addr / d->mb_w;
    t.cur = &d->mb[addr];
    memset(t.cur, 0, sizeof *t.cur);
    t.cur->slice = t.slice_id;
    int skip = 0;
    if (!islice) skip = read_mb_skip(&t);
    int rc;
    if (skip) {
      rc = t.bslice ? decode_b_skip(&t) : decode_skip(&t);
      t.cur->qp = t.qp;
    } else {
      rc = decode_mb(&t, islice);
      if (t.cur->kind == K_PCM) t.cur->qp = t.qp;
    }
    if (rc) return rc;
    finish_info(&t);
    if (terminate(&t.c)) break;
    addr++;
  }
  return OD_OK;
}

/* ---- pictures ------------------------------------------------------------------------------------- */

static int finish_picture(od_dec *d) {
  if (!d->have_cur) return OD_OK;
  const sps_t *s = &d->sps[d->cur_sps];
  const int nmb = d->mb_w * d->mb_h;
  for (int i = 0; i < nmb; i++)
    if (d->mb[i].slice == 0) {
      free_cur(d);
      return fail(d, OD_EBITSTREAM, "picture with undecoded macroblocks");
    }
  if (d->nframes >= OD_MAX_FRAMES) return fail(d, OD_EUNSUPPORTED, "too many pictures");
  frame_t *f = &d->frames[d->nframes++];
  memset(f, 0, sizeof *f);
  f->used = 1;
  f->id = ++d->next_id;
  f->frame_num = d->cur_hdr.frame_num;
  f->poc = d->cur_poc;
  f->idr_seq = d->idr_seq;
  f->mb_w = d->mb_w;
  f->mb_h = d->mb_h;
  memcpy(f->crop, s->crop, sizeof f->crop);
  const size_t yl = (size_t)nmb * 256, cl = (size_t)nmb * 64;
  f->y = (uint8_t *)malloc(yl);
  f->cb = (uint8_t *)malloc(cl);
  f->cr = (uint8_t *)malloc(cl);
  f->y16 = (uint16_t *)malloc(yl * 2);
  f->cb16 = (uint16_t *)malloc(cl * 2);
  f->cr16 = (uint16_t *)malloc(cl * 2);
  f->mbs = (ia_mbinfo *)malloc(nmb * sizeof(ia_mbinfo));
  if (!f->y || !f->cb || !f->cr || !f->y16 || !f->cb16 || !f->cr16 || !f->mbs) return fail(d, OD_ENOMEM, "memory");
  for (size_t i = 0; i < yl; i++) f->y[i] = (uint8_t)d->py[i];
  for (size_t i = 0; i < cl; i++) {
    f->cb[i] = (uint8_t)d->pcb[i];
    f->cr[i] = (uint8_t)d->pcr[i];
  }
  memcpy(f->mbs, d->info, nmb * sizeof(ia_mbinfo));
  or_deblock_picture_params(f->y, f->cb, f->cr, f->mbs, d->mb_w, d->mb_h, d->sfilter);
  for (size_t i = 0; i < yl; i++) f->y16[i] = f->y[i];
  for (size_t i = 0; i < cl; i++) {
    f->cb16[i] = f->cb[i];
    f->cr16[i] = f->cr[i];
  }
  mark(d, s, &d->cur_hdr, f);
  free_cur(d);
  return OD_OK;
}

static int start_picture(od_dec *d, const sh_t *h) {
  const pps_t *q = &d->pps[h->pps_id];
  const sps_t *s = &d->sps[q->sps_id];
  d->have_cur = 1;
  d->cur_sps = q->sps_id;
  d->cur_pps = h->pps_id;
  d->cur_hdr = *h;
  d->mb_w = s->mb_w;
  d->mb_h = s->mb_h;
  d->nslices = 0;
  const size_t nmb = (size_t)s->mb_w * s->mb_h;
  d->py = (uint16_t *)calloc(nmb * 256, 2);
  d->pcb = (uint16_t *)calloc(nmb * 64, 2);
  d->pcr = (uint16_t *)calloc(nmb * 64, 2);
  d->mb = (dmb_t *)calloc(nmb, sizeof(dmb_t));
  d->info = (ia_mbinfo *)calloc(nmb, sizeof(ia_mbinfo));
  if (!d->py || !d->pcb || !d->pcr || !d->mb || !d->info) return fail(d, OD_ENOMEM, "memory");
  if (h->idr) {
    d->idr_seq++;
    for (int i = 0; i < d->nframes; i++) d->frames[i].is_ref = 0;
  }
  d->cur_poc = compute_poc(d, s, h);
  return OD_OK;
}

int od_decode_nal(od_dec *d, const uint8_t *nal, size_t len) {
  if (!d || !nal || len < 1) return OD_EINVAL;
  const int ref_idc = (nal[0] >> 5) & 3, type = nal[0] & 31;
  uint8_t *rbsp = (uint8_t *)malloc(len);
  if (!rbsp) return fail(d, OD_ENOMEM, "memory");
  const size_t n = unescape(nal + 1, len - 1, rbsp);
  br_t b = {rbsp, n, 0};
  int rc = OD_OK;
  if (type == 7) rc = parse_sps(d, &b);
  else if (type == 8) rc = parse_pps(d, &b);
  else if (type == 1 || type == 5) {
    sh_t h;
    rc = parse_slice_header(d, &b, type, ref_idc, &h);
    if (!rc) {
      /* a slice starting at macroblock 0 starts a picture (slices arrive in order) */
      if (h.first_mb == 0 || !d->have_cur) {
        rc = finish_picture(d);
        if (!rc) rc = start_picture(d, &h);
      }
      if (!rc) rc = decode_slice(d, &h, &b);
    }
  }
  free(rbsp);
  return rc;
}

int od_decode_annexb(od_dec *d, const uint8_t *data, size_t len) {
  if (!d || !data) return OD_EINVAL;
  /* NAL units follow 00 00 01 start codes; trailing zero bytes belong to the next start code */
  size_t i = 0, start = 0;
  int in_nal = 0;
  while (i + 2 < len) {
    if (data[i] == 0 && data[i + 1] == 0 && data[i + 2] == 1) {
      if (in_nal) {
        size_t e = i;
        while (e > start && data[e - 1] == 0) e--;
        if (e > start) {
          const int rc = od_decode_nal(d, data + start, e - start);
          if (rc) return rc;
        }
      }
      i += 3;
      start = i;
      in_nal = 1;
    } else {
      i++;
    }
  }
  if (in_nal) {
    size_t e = len;
    while (e > start && data[e - 1] == 0) e--;
    if (e > start) return od_decode_nal(d, data + start, e - start);
  }
  return OD_OK;
}

int od_flush(od_dec *d) {
  if (!d) return OD_EINVAL;
  int rc = finish_picture(d);
  if (rc) return rc;
  /* output order: by IDR period, then order count (stable for equal counts) */
  for (int i = 0; i < d->nframes; i++) d->frames[i].order = i;
  for (int a = 0; a < d->nframes; a++)
    for (int b = a + 1; b < d->nframes; b++) {
      frame_t *x = &d->frames[a], *y = &d->frames[b];
      if (y->idr_seq < x->idr_seq || (y->idr_seq == x->idr_seq && (y->poc < x->poc || (y->poc == x->poc && y->order < x->order)))) {
        frame_t tmp = *x;
        *x = *y;
        *y = tmp;
      }
    }
  return OD_OK;
}

int od_frame_count(const od_dec *d) { return d ? d->nframes : 0; }

int od_frame_size(const od_dec *d, int k, int *width, int *height, int *mb_w, int *mb_h) {
  if (!d || k < 0 || k >= d->nframes) return OD_EINVAL;
  const frame_t *f = &d->frames[k];
  if (width) *width = f->mb_w * 16 - f->crop[0] - f->crop[1];
  if (height) *height = f->mb_h * 16 - f->crop[2] - f->crop[3];
  if (mb_w) *mb_w = f->mb_w;
  if (mb_h) *mb_h = f->mb_h;
  return OD_OK;
}

int od_frame_read(const od_dec *d, int k, uint8_t *y, uint8_t *cb, uint8_t *cr) {
  if (!d || k < 0 || k >= d->nframes) return OD_EINVAL;
  const frame_t *f = &d->frames[k];
  const int w = f->mb_w * 16 - f->crop[0] - f->crop[1], h = f->mb_h * 16 - f->crop[2] - f->crop[3];
  const int pw = f->mb_w * 16, cw = (w + 1) / 2, ch = (h + 1) / 2;
  for (int r = 0; r < h; r++) memcpy(y + (size_t)r * w, f->y + (size_t)(r + f->crop[2]) * pw + f->crop[0], w);
  for (int r = 0; r < ch; r++) {
    memcpy(cb + (size_t)r * cw, f->cb + (size_t)(r + f->crop[2] / 2) * (pw / 2) + f->crop[0] / 2, cw);
    memcpy(cr + (size_t)r * cw, f->cr + (size_t)(r + f->crop[2] / 2) * (pw / 2) + f->crop[0] / 2, cw);
  }
  return OD_OK;
}

int od_frame_read_full(const od_dec *d, int k, uint8_t *y, uint8_t *cb, uint8_t *cr, ia_mbinfo *mbs) {
  if (!d || k < 0 || k >= d->nframes) return OD_EINVAL;
  const frame_t *f = &d->frames[k];
  const size_t nmb = (size_t)f->mb_w * f->mb_h;
  if (y) memcpy(y, f->y, nmb * 256);
  if (cb) memcpy(cb, f->cb, nmb * 64);
  if (cr) memcpy(cr, f->cr, nmb * 64);
  if (mbs) memcpy(mbs, f->mbs, nmb * sizeof(ia_mbinfo));
  return OD_OK;
}

int od_cabac_decode_bins(const uint8_t *data, size_t n, const uint8_t *kinds, const uint16_t *ctx, int count, int intra,
                         int qp, uint8_t *vals, int *final_terminate, size_t *bytes_read) {
  if (!data || !kinds || !ctx || !vals || count < 0) return OD_EINVAL;
  br_t b = {data, n, 0};
  cabac_t c;
  memset(&c, 0, sizeof c);
  c.b = &b;
  cabac_init_contexts(&c, intra, 0, qp);
  cabac_init_engine(&c);
  for (int i = 0; i < count; i++) {
    if (kinds[i] == 0) {
      if (ctx[i] >= 1024) return OD_EINVAL;
      vals[i] = (uint8_t)decision(&c, ctx[i]);
    } else if (kinds[i] == 1) {
      vals[i] = (uint8_t)bypass(&c);
    } else {
      vals[i] = (uint8_t)terminate(&c);
    }
  }
  if (final_terminate) *final_terminate = terminate(&c);
  if (bytes_read) *bytes_read = (b.pos + 7) / 8;
  return OD_OK;
}

int od_stats(const od_dec *d, int64_t *out, int n) {
  if (!d || !out || n < 0) return OD_EINVAL;
  for (int i = 0; i < n && i < OD_NSTATS; i++) out[i] = d->stats[i];
  return OD_NSTATS;
}
