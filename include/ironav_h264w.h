/* ironav_h264w.h — C ABI of the host side around the GPU analysis: the encoder driver's picture-type
 * decisions and the bitstream back end that writes the analysed macroblocks.
 *
 * The GPU analysis (ironav_h264a.h) returns one `ia_mbcode` per macroblock. What nervosys/IronAV's
 * encoder does around its slice coders stays on the CPU, as in the reference:
 *  - the picture type of each input picture: first picture, keyint, forced, or a scene cut
 *    (`H264Encoder::encode`, crates/ironav-codec/src/h264/enc/mod.rs:700-717, with `change_threads`,
 *    crates/ironav-codec/src/change.rs:57-95);
 *  - the bitstream: SPS, PPS, slice headers and CABAC slice data written from the macroblock codes
 *    (`write_mb_cabac` and `mb_bins`, enc/frame.rs:2020-2279; `CabacWriter`, enc/cabac.rs:42-146;
 *    `BitWriter` and NAL escaping, enc/bitw.rs; `sps`/`pps`/`slice_header`, enc/mod.rs:296-400).
 *
 * Conventions as in ironav_h264a.h: every function returns IW_OK (0) or a negative IW_E* code; the
 * caller owns every buffer; an object is not thread safe. Pictures are 8-bit 4:2:0 padded to whole
 * macroblocks (luma stride mb_w*16).
 */
#ifndef IRONAV_H264W_H
#define IRONAV_H264W_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#include "ironav_h264a.h"

enum {
  IW_OK = 0,
  IW_EINVAL = -1,        /* an argument is out of range */
  IW_ENOMEM = -3,        /* an allocation failed */
  IW_EUNSUPPORTED = -5,  /* a configuration the back end does not write (CAVLC) */
  IW_ESPACE = -6         /* the output buffer is too small */
};

/* ---- picture types (enc/mod.rs:700-717, change.rs) ---- */

/* The low-resolution luma of a padded picture: the mean of each 4x4 block (rounded down), (mb_w*4) x
 * (mb_h*4) values (`LowRes::new` as `H264Encoder::low_res` calls it, enc/mod.rs:686-692,
 * change.rs:18-25). */
int iw_lowres(const uint8_t *y, int mb_w, int mb_h, int32_t *out);

/* How well `cur` is predicted from `prev` (two low-resolution pictures lw x lh): `inter` sums each
 * 8x8 block's least SAD against `prev` over displacements within +-4, `intra` its absolute error
 * against its own mean (`change_threads`, change.rs:57-89); the same sums for any `threads`. */
int iw_change(const int32_t *prev, const int32_t *cur, int lw, int lh, int threads, int64_t *inter,
              int64_t *intra);

typedef struct iw_gop iw_gop;

/* The picture-type state of one encoder: picture count, last IDR picture and the previous picture's
 * low-resolution luma. keyint 0 takes the reference's default (250, enc/mod.rs:27). */
int iw_gop_create(iw_gop **out, int width, int height, uint64_t keyint, int threads);
void iw_gop_destroy(iw_gop *g);

/* The type of the next input picture (padded luma y): *kind = 0 (IDR: first picture, keyint reached,
 * `force_idr`, or a scene cut: inter > 0.8 x intra) or 1 (P). inter/intra (may be NULL) receive the
 * scene-cut measure against the previous picture (0 for the first). */
int iw_gop_next(iw_gop *g, const uint8_t *y, int force_idr, int *kind, int64_t *inter, int64_t *intra);

/* ---- the bitstream (enc/mod.rs:296-400, :539-652; enc/frame.rs:2020-2339; enc/cabac.rs; enc/bitw.rs) ---- */

typedef struct iw_params {
  int32_t width, height;      /* luma size in samples (even); cropped from whole macroblocks in the SPS */
  int32_t fps_num, fps_den;   /* frame rate for the level and the VUI timing; 0/0 is 25 (enc/mod.rs:291-294) */
  int32_t cabac;              /* 1 (CABAC; the CAVLC writer is not part of this back end) */
  int32_t transform8;         /* transform_8x8_mode_flag (High profile) */
  int32_t nrefs;              /* references P pictures use: max_num_ref_frames */
  int32_t slices;             /* slice count: the bands i*mb_h/n .. (i+1)*mb_h/n (enc/mod.rs:603) */
  int32_t bframes;            /* 0: IPPP with order counts of type 2; > 0: B pictures may be written
                                 (iw_write_picture_at), order counts of type 0 in the slice headers and
                                 one picture of reordering (enc/mod.rs:322-327, 370-372) */
} iw_params;

typedef struct iw_stream iw_stream;

/* One output stream, coded IPPP as `code_picture` codes it without B pictures: SPS (High, or Main
 * without the 8x8 transform; order counts of type 2; VUI timing), PPS, then per picture one slice NAL
 * unit per band. */
int iw_stream_create(iw_stream **out, const iw_params *p);
void iw_stream_destroy(iw_stream *s);

/* Writes the access unit of one picture analysed by the GPU: kind 0 (IDR: SPS, PPS and IDR slices) or
 * 1 (P slices against the `nref` most recent references), slice QP `qp`, `codes` the ia_mbcode of every
 * macroblock in raster order (ia_read_analysis). *len receives its size in bytes; iw_read_picture
 * copies it. The macroblocks' CABAC bins and contexts are those the analysis counted its rate on. */
int iw_write_picture(iw_stream *s, int kind, int qp, int nref, const ia_mbcode *codes, size_t *len);
int iw_read_picture(const iw_stream *s, uint8_t *out, size_t cap);

/* The same for a stream with bframes > 0, in coding order, each picture with its display index (the
 * reference's `d`, enc/mod.rs:539-547): kind 2 writes a B picture (B slices with spatial direct
 * prediction against one list 0 and one list 1 picture, not kept as a reference: nal_ref_idc 0, no
 * frame_num step; enc/mod.rs:566-600, 636-641); its picture order count is 2 x (display - the last IDR
 * picture's display). */
int iw_write_picture_at(iw_stream *s, int kind, int qp, int nref, int64_t display, const ia_mbcode *codes,
                        size_t *len);

/* level_idc for a picture size and frame rate: the first level of Table A-1 whose frame size and
 * macroblock rate hold it, else the first whose frame size does (enc/mod.rs:155-158). */
int iw_level(int width, int height, int fps_num, int fps_den);

/* A NAL unit with a four-byte start code and emulation prevention bytes (`put_nal`, enc/bitw.rs:93-105);
 * with out NULL only *len is computed. */
int iw_put_nal(int ref_idc, int type, const uint8_t *rbsp, size_t n, uint8_t *out, size_t cap, size_t *len);

/* The CABAC coder on its own (for checking it against a decoder, as enc/cabac.rs:252-290 does): bins
 * kinds[i] (0 decision on context ctx[i], 1 bypass, 2 a terminate bin of 0) with values vals[i], from
 * contexts initialised as an I (intra) or P slice's at `qp`, ended by a terminate bin of 1. */
int iw_cabac_encode_bins(const uint8_t *kinds, const uint16_t *ctx, const uint8_t *vals, int n, int intra, int qp,
                         uint8_t *out, size_t cap, size_t *len);

#ifdef __cplusplus
}
#endif
#endif
