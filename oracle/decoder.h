/* decoder.h — the verifier: a C restatement of nervosys/IronAV's H.264 decoder for the subset the
 * encoder's streams use.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (ironav_amd/) includes, links or calls this code;
 * only tests/ load it, to check that bitstreams decode to the pictures the
 * encoder reconstructed (crates/ironav-codec/src/h264/enc/mod.rs:812-825 is the reference's test of
 * the same property) and to pin the oracle's reconstruction primitives against the reference's
 * decoder fixtures (libx264 streams with FFmpeg's decoded pictures,
 * crates/ironav-engine/tests/data/h264_*, crates/ironav-engine/tests/e2e.rs:1991-2036).
 *
 * Restated from the reference decoder: crates/ironav-codec/src/h264/mod.rs (NAL dispatch, picture
 * boundaries), ps.rs (SPS, PPS, slice header), bits.rs (emulation prevention, Exp-Golomb),
 * cabac.rs (the arithmetic decoder), mb.rs (CABAC macroblock layer, motion vector prediction,
 * residual, intra and inter reconstruction), dpb.rs (order counts, reference lists, marking),
 * deblock.rs (through the oracle's loop filter). The subset: progressive frames, 8-bit 4:2:0,
 * I, P and B slices, CABAC, the 8x8 transform, scaling matrices, explicit and implicit weighted
 * prediction, spatial and temporal direct prediction, I_PCM, several slices and reference pictures.
 * Anything else (CAVLC, MBAFF, field pictures, other chroma formats or bit depths) is rejected with
 * OD_EUNSUPPORTED.
 */
#ifndef IRONAV_ORACLE_DECODER_H
#define IRONAV_ORACLE_DECODER_H

#include <stddef.h>
#include <stdint.h>

#include "../include/ironav_h264a.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { OD_OK = 0, OD_EINVAL = -1, OD_EBITSTREAM = -2, OD_EUNSUPPORTED = -3, OD_ENOMEM = -4 };

typedef struct od_dec od_dec;

od_dec *od_create(void);
void od_destroy(od_dec *d);
/* The text of the last error. */
const char *od_error(const od_dec *d);

/* One NAL unit (header byte first, emulation prevention bytes still in). */
int od_decode_nal(od_dec *d, const uint8_t *nal, size_t len);
/* A byte stream of NAL units with start codes (Annex B). */
int od_decode_annexb(od_dec *d, const uint8_t *data, size_t len);
/* Finishes the picture being decoded; after it every decoded picture is available. */
int od_flush(od_dec *d);

/* Decoded pictures, in output order (order count within each IDR period) after od_flush. */
int od_frame_count(const od_dec *d);
/* Cropped size of frame k and its size in macroblocks. */
int od_frame_size(const od_dec *d, int k, int *width, int *height, int *mb_w, int *mb_h);
/* The cropped planes of frame k (y: width x height, cb/cr: ceil(width/2) x ceil(height/2)). */
int od_frame_read(const od_dec *d, int k, uint8_t *y, uint8_t *cb, uint8_t *cr);
/* The whole planes of frame k (mb_w*16 x mb_h*16 luma, half that chroma), loop-filtered, and its
 * macroblock records (mb_w*mb_h; kind, qp, cbp, skip, t8x8, nnz, vectors and references); any
 * pointer may be NULL. */
int od_frame_read_full(const od_dec *d, int k, uint8_t *y, uint8_t *cb, uint8_t *cr, ia_mbinfo *mbs);

/* What the decoded syntax exercised, as counts: Intra4x4 modes 0-8, Intra8x8 modes 0-8, Intra16x16
 * modes 0-3, chroma modes 0-3, P partitionings 0-3 (16x16, 16x8, 8x16, 8x8), sub-partitionings 0-3,
 * P_Skip, I_PCM, inter macroblocks with the 8x8 transform, reference indexes above 0. Returns
 * OD_NSTATS and fills min(n, OD_NSTATS) counts. */
enum {
  OD_STAT_I4 = 0, OD_STAT_I8 = 9, OD_STAT_I16 = 18, OD_STAT_CHROMA = 22, OD_STAT_PART = 26, OD_STAT_SUB = 30,
  OD_STAT_SKIP = 34, OD_STAT_PCM = 35, OD_STAT_INTER_T8 = 36, OD_STAT_REF_GT0 = 37,
  /* B slices: B_Skip, B_Direct_16x16, B 16x16 from list 0 / list 1 / both, B 16x8 / 8x16, B_8x8, its
     13 sub_mb_types, macroblocks with temporal direct prediction, bi-predicted blocks with implicit and
     with explicit weights */
  OD_STAT_B_SKIP = 38, OD_STAT_B_DIRECT16 = 39, OD_STAT_B16 = 40, OD_STAT_B_PART = 43, OD_STAT_B8X8 = 45,
  OD_STAT_B_SUB = 46, OD_STAT_B_TEMPORAL = 59, OD_STAT_B_IMPLICIT = 60, OD_STAT_B_EXPLICIT = 61, OD_NSTATS = 62
};
int od_stats(const od_dec *d, int64_t *out, int n);

/* The arithmetic decoder on its own: `count` bins of kinds[i] (0 decision on context ctx[i], 1 bypass,
 * 2 terminate) from contexts initialised as an I (intra) or P (cabac_init_idc 0) slice's at `qp`; then
 * one more terminate bin into *final_terminate and the bytes consumed into *bytes_read. */
int od_cabac_decode_bins(const uint8_t *data, size_t n, const uint8_t *kinds, const uint16_t *ctx, int count, int intra,
                         int qp, uint8_t *vals, int *final_terminate, size_t *bytes_read);

#ifdef __cplusplus
}
#endif
#endif
