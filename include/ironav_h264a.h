/* ironav_h264a.h — C ABI of the MI355X (gfx950) H.264 per-macroblock analysis stage.
 *
 * The stage replaces the per-slice analysis loop of nervosys/IronAV's software H.264 encoder:
 * the `FrameCoder::encode` loop that calls `code_mb` for every macroblock of a slice
 * (crates/ironav-codec/src/h264/enc/frame.rs:452-487, code_mb at :715) together with the half-sample
 * planes of each reference picture (`HalfPel::new`, crates/ironav-codec/src/h264/enc/subpel.rs:42-115).
 * The caller keeps what stays on the CPU in the reference encoder: picture types and QP
 * (enc/mod.rs:404-416), CABAC/CAVLC writing of the returned macroblock codes (enc/frame.rs:1922-2279),
 * deblocking (h264/deblock.rs:410) and the reference list (enc/mod.rs:559,663-667).
 *
 * Conventions
 *  - Every function returns IA_OK (0) or a negative IA_E* code; nothing unwinds across the ABI.
 *    ia_last_error() describes the last failure of a context.
 *  - The caller owns every host buffer. The context owns its device buffers and `nrefs` reference slots.
 *  - Pictures are 8-bit 4:2:0, padded to whole macroblocks by repeating the right column and bottom
 *    row (the reference's `H264Encoder::source`, enc/mod.rs:481-529): luma is mb_w*16 x mb_h*16 with
 *    stride mb_w*16, each chroma plane mb_w*8 x mb_h*8 with stride mb_w*8.
 *  - Within each slice the results are those of running `code_mb` in raster order; slices are the
 *    bands i*mb_h/n .. (i+1)*mb_h/n of enc/mod.rs:603 with n = ia_config.slices.
 *  - A context is not thread safe; use one per encoder thread.
 */
#ifndef IRONAV_H264A_H
#define IRONAV_H264A_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  IA_OK = 0,
  IA_EINVAL = -1,        /* an argument is out of range */
  IA_EDEVICE = -2,       /* a HIP call failed, or no gfx950 device is present */
  IA_ENOMEM = -3,        /* a device allocation failed */
  IA_ESTATE = -4,        /* the call does not fit the context's state (e.g. an empty reference slot) */
  IA_EUNSUPPORTED = -5   /* a configuration the GPU path does not implement */
};

/* Macroblock kinds, numbered as h264/mb.rs:21-30 `Kind`. */
enum { IA_KIND_NONE = 0, IA_KIND_I4X4 = 1, IA_KIND_I8X8 = 2, IA_KIND_I16X16 = 3, IA_KIND_IPCM = 4, IA_KIND_INTER = 5 };

/* Integer motion search of `motion_search` (enc/frame.rs:949-1019):
 *  IA_SEARCH_DIAMOND  the reference's predictor candidates and diamond, exactly;
 *  IA_SEARCH_FULL     the same with one more candidate per partition and reference: the exhaustive
 *                     integer-sample SAD winner over +-search_range (see ia_motion_search);
 *  IA_SEARCH_LEARNED  the same with the learned proposal's winner as that candidate (see
 *                     ia_load_proposal_weights). */
enum { IA_SEARCH_DIAMOND = 0, IA_SEARCH_FULL = 1, IA_SEARCH_LEARNED = 2 };

/* Picture kinds of `code_picture` (enc/mod.rs:539): 0 IDR (I slices), 1 P, 2 B (ia_analyse_b). */
enum { IA_PIC_IDR = 0, IA_PIC_P = 1, IA_PIC_B = 2 };

typedef struct ia_config {
  int32_t width, height;   /* luma size in samples (even); mb_w = ceil(width/16), mb_h = ceil(height/16) */
  int32_t bit_depth;       /* 8 */
  int32_t cabac;           /* 1: High profile, decisions by rate-distortion with the CABAC bit model
                              (enc/frame.rs:757-758); 0: decisions by SATD (CAVLC) */
  int32_t transform8;      /* transform_8x8_mode_flag; used with cabac only (enc/frame.rs:409) */
  int32_t sub8x8;          /* P_8x8 sub-partitions, active at qp <= 20 (enc/frame.rs:200,436) */
  int32_t fast;            /* `realtime()` decisions (enc/frame.rs:348): shorter search, 16x16 only, early P_Skip */
  int32_t nrefs;           /* reference slots, 1..16 */
  int32_t slices;          /* slice count, 1..mb_h (fixed: it changes the decisions) */
  int32_t search_mode;     /* IA_SEARCH_* */
  int32_t search_range;    /* integer-sample range of the full search, 1..64 */
  float fallback_ratio;    /* learned proposal: fall back to full search when cost > ratio x predicted;
                              a non-finite ratio (infinity) turns the fallback off */
} ia_config;

/* What later pictures and the loop filter read about a coded macroblock: the frame-macroblock subset
 * of h264/mb.rs:34-68 `MbInfo` that `finish_mb` (enc/frame.rs:2281-2339) fills. List 1 is used by B
 * pictures only; P and I pictures leave it at vector 0, reference -1. */
typedef struct ia_mbinfo {
  uint16_t slice;          /* 1-based slice id */
  uint8_t kind;            /* IA_KIND_* */
  int8_t qp;
  uint8_t cbp;
  uint8_t skip;
  uint8_t t8x8;
  uint8_t reserved;
  uint8_t nnz[16];         /* luma coefficient counts, raster 4x4 blocks */
  int16_t mv[16][2];       /* list 0 vectors per raster 4x4 block, quarter samples */
  int8_t ref_idx[4];       /* list 0 reference index per 8x8 block, -1 when unused */
  int32_t ref_id[4];       /* 1 + ref_idx, -1 when unused (B pictures: 1, the list 0 picture) */
  int16_t mv1[16][2];      /* list 1 vectors per raster 4x4 block (B pictures) */
  int8_t ref_idx1[4];      /* list 1 reference index per 8x8 block, -1 when unused */
  int32_t ref_id1[4];      /* 2 (the list 1 picture) where list 1 is used, else -1 */
  uint8_t bdirect;         /* B_Skip or B_Direct_16x16 */
  uint8_t direct8;         /* B_8x8: mask of the 8x8 blocks coded B_Direct_8x8 (15 for bdirect) */
  uint8_t reserved1[2];
} ia_mbinfo;

/* A macroblock's decisions and levels, ready to write: `MbCode` (enc/frame.rs:275-310). Levels fit
 * int16 at 8 bits. In B pictures `part` is 0..3 (B_8x8 has no sub-partitions) and the motion is in
 * bdir/buse/bmv/bmvd; `skip` marks B_Skip, and part 0 with bdir[0] 0 B_Direct_16x16. */
typedef struct ia_mbcode {
  uint8_t kind;            /* IA_KIND_* */
  uint8_t skip;            /* P_Skip */
  uint8_t i16_mode;
  uint8_t chroma_mode;
  uint8_t part;            /* 0 16x16, 1 16x8, 2 8x16, 3 8x8 */
  uint8_t t8x8;
  uint8_t cbp;
  uint8_t reserved;
  uint8_t modes[16];       /* Intra4x4/8x8 modes, raster 4x4 blocks */
  uint8_t refs[4];         /* reference index per partition (per 8x8 block for P_8x8) */
  uint8_t sub[4];          /* sub_mb_type per 8x8 block of P_8x8 */
  uint8_t nnz[16];
  uint8_t cnnz[2][4];
  int16_t mvs[16][2];      /* per partition in coding order (sub-partitions of P_8x8 in coding order) */
  int16_t mvds[16][2];     /* vector minus its predictor, per partition */
  int16_t luma[16][16];    /* per raster 4x4 block, zig-zag scan order (DC slot unused for Intra16x16) */
  int16_t luma8[4][64];    /* per 8x8 block, 8x8 zig-zag scan order, when t8x8 */
  int16_t dc[16];          /* Intra16x16 luma DC, scan order */
  int16_t cdc[2][4];       /* chroma DC per component */
  int16_t cac[2][4][16];   /* chroma AC per component and 4x4 block, scan order (slot 0 unused) */
  uint8_t bdir[4];         /* B: per partition 0 direct, 1 list 0, 2 list 1, 3 both */
  uint8_t buse[4];         /* B: per 8x8 block, bit 0 list 0 used, bit 1 list 1 used */
  int16_t bmv[2][16][2];   /* B: per list, the vector of each raster 4x4 block */
  int16_t bmvd[2][4][2];   /* B: per list, each partition's vector difference */
} ia_mbcode;

/* The costs the decisions were taken on, for checking and for rate studies. */
typedef struct ia_mbstats {
  int32_t skip_sad;        /* SAD of the P_Skip prediction (enc/frame.rs:729); -1 in I slices */
  int32_t i16_cost;        /* best Intra16x16 SATD + lambda cost (enc/frame.rs:812) */
  int32_t inter_cost;      /* best_inter's cost (enc/frame.rs:839); -1 when not searched */
  int32_t inter_part;      /* best_inter's partitioning; -1 when not searched */
  double rd_cost;          /* J of the kept candidate (enc/frame.rs:1453-1457); 0 when decided before RD */
  int32_t bits;            /* CABAC: the kept macroblock's rate in 1/256 bits, counted on the live contexts
                              as the writer codes it (enc/frame.rs:2020-2024, enc/cabac.rs:150-191); 0 with
                              CAVLC */
} ia_mbstats;

/* One exhaustive-search winner: integer-sample vector in quarter samples, and its SAD. */
typedef struct ia_fsresult {
  int16_t mv[2];
  uint32_t sad;
} ia_fsresult;

/* Partitions of a macroblock searched by ia_motion_search, 41 per macroblock and reference:
 *   0        16x16
 *   1..2     16x8  (top, bottom)
 *   3..4     8x16  (left, right)
 *   5..8     8x8   (raster)
 *   9..16    8x4   (9 + 2*b8 + half)
 *   17..24   4x8   (17 + 2*b8 + half)
 *   25..40   4x4   (25 + 4*b8 + raster index inside the 8x8 block) */
#define IA_NPART 41

typedef struct ia_ctx ia_ctx;

/* Creates a context on the current HIP device: checks the configuration, allocates the device
 * buffers of one picture, its `nrefs` reference slots and the scratch of the analysis. Replaces
 * the per-picture set-up of `FrameCoder::new` (enc/frame.rs:381-413). */
int ia_create(ia_ctx **out, const ia_config *cfg);

/* Frees every device buffer of the context. */
void ia_destroy(ia_ctx *ctx);

/* The last failure of `ctx` (or of ia_create when ctx is NULL), as text. */
const char *ia_last_error(const ia_ctx *ctx);

/* Enqueues the context's work on `hip_stream` (a hipStream_t; NULL for the context's own stream). */
int ia_set_stream(ia_ctx *ctx, void *hip_stream);

/* Loads a reference picture into `slot`: its deblocked luma and chroma (host, padded layout) and
 * its macroblock records (mb_w*mb_h, for the co-located vector of enc/frame.rs:944-947; may be NULL
 * for none inter-coded), then builds its half-sample planes on the GPU (`HalfPel::new`,
 * enc/subpel.rs:57-115). Synchronous. */
int ia_set_reference(ia_ctx *ctx, int slot, const uint8_t *y, const uint8_t *cb, const uint8_t *cr,
                     const ia_mbinfo *mbs);

/* Copies the 4 half-sample planes of `slot` (G, B, H, J; each (mb_w*16+96) x (mb_h*16+96), 48 samples
 * of border) to `out`. For checking. */
int ia_read_halfpel(ia_ctx *ctx, int slot, uint8_t *out);

/* Makes the last analysed picture the reference in `slot` without a host round trip: its
 * reconstruction and macroblock records are copied on the device; with `deblock` non-zero the copy
 * is loop-filtered first, as the encoder filters each picture before keeping it
 * (enc/mod.rs:653-654, h264/deblock.rs:608-695: frame macroblocks, 4:2:0, filtering across slice
 * edges, zero offsets); then its half-sample planes are built (enc/mod.rs:660). Replaces the
 * reference's deblock_picture + HalfPel::new + reference list push (enc/mod.rs:653-667). The
 * analysis outputs (ia_read_analysis) stay the reconstruction before filtering. Synchronous. */
int ia_promote_reference(ia_ctx *ctx, int slot, int deblock);

/* Copies reference `slot`: its luma (mb_w*16 x mb_h*16), chroma and records; any pointer may be
 * NULL. Synchronous. */
int ia_read_reference(ia_ctx *ctx, int slot, uint8_t *y, uint8_t *cb, uint8_t *cr, ia_mbinfo *mbs);

/* Uploads the picture to analyse (host, padded layout) to the device, on the context's stream. */
int ia_upload_source(ia_ctx *ctx, const uint8_t *y, const uint8_t *cb, const uint8_t *cr);

/* Exhaustive integer-sample SAD search of the uploaded picture against the references in
 * `ref_slots[0..nref)`, for all 41 partitions of every macroblock. The window of macroblock
 * (mbx, mby) is dx in [max(-R, -(16*mbx+24)), min(R, 16*mb_w-16*mbx+8)], and likewise dy, R the
 * search range: the vectors `motion_search` accepts for the whole macroblock (enc/frame.rs:952-954).
 * The winner is the least SAD, ties to the first in raster order of (dy, dx). Results stay on the
 * device for ia_analyse_uploaded and ia_read_motion. Asynchronous on the context's stream; this is
 * the throughput kernel of the stage. */
int ia_motion_search(ia_ctx *ctx, const int *ref_slots, int nref);

/* Copies the last ia_motion_search results, laid out [ref][mb][IA_NPART]. Synchronous. */
int ia_read_motion(ia_ctx *ctx, ia_fsresult *out);

/* Analyses the uploaded picture: `code_mb` for every macroblock of every slice
 * (enc/frame.rs:715-790), with the CABAC context states carried from one macroblock to the next as
 * the writer carries them (enc/frame.rs:465,2020-2024). kind is IA_PIC_IDR or IA_PIC_P; qp the
 * slice QP (IDR pictures are coded at qp-3 by the reference's driver, enc/mod.rs:36,404).
 * With IA_SEARCH_FULL the last ia_motion_search results over the same ref_slots are used.
 * Asynchronous on the context's stream. */
int ia_analyse_uploaded(ia_ctx *ctx, int qp, int kind, const int *ref_slots, int nref);

/* Analyses the uploaded picture as a B picture: `decide_b` for every macroblock of every slice
 * (enc/frame.rs:1518-1780: B_Skip, B_Direct_16x16, list 0 / list 1 / bi-predicted 16x16, 16x8, 8x16
 * and 8x8 partitionings with B_Direct_8x8 blocks, Intra16x16, Intra4x4, by rate-distortion cost with
 * the CABAC bit model, spatial direct prediction). Replaces the slice coders of `code_picture` for a
 * picture of kind 2 built `.with_b(l1, dist)` (enc/mod.rs:551-566, 619-621, enc/frame.rs:442-447):
 * l0_slot holds the nearest reference shown before the picture, l1_slot the nearest shown after it
 * (its records give the co-located motion), dist_l0 and dist_l1 their picture-order distances from
 * the list 0 picture to this one and to the list 1 picture (0 < dist_l0 < dist_l1). CABAC only (the
 * reference codes B pictures only with CABAC, enc/mod.rs:534-536). The search mode's exhaustive or
 * learned candidates apply to P pictures; B pictures use the reference's candidates. A B picture is
 * not a reference: ia_promote_reference refuses it. The context needs nrefs >= 2 (the reference keeps
 * two references when it codes B pictures, enc/mod.rs:287-289). Asynchronous on the context's stream. */
int ia_analyse_b(ia_ctx *ctx, int qp, int l0_slot, int l1_slot, int dist_l0, int dist_l1);

/* Copies the results of the last analysis. Any output pointer may be NULL. Synchronous. */
int ia_read_analysis(ia_ctx *ctx, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y, uint8_t *recon_cb,
                     uint8_t *recon_cr, ia_mbinfo *mbs);

/* Upload, (full search,) analysis and read-back in one call: the drop-in for one picture's
 * `FrameCoder::new(..).with_refs(..).encode(..)` across all slices, minus the entropy coding. */
int ia_analyse_picture(ia_ctx *ctx, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp, int kind,
                       const int *ref_slots, int nref, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y,
                       uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs);

/* Waits for the context's stream. */
int ia_sync(ia_ctx *ctx);

/* The learned motion proposal (IA_SEARCH_LEARNED; no counterpart in the reference). With it,
 * ia_motion_search proposes 8 candidate cells per macroblock and reference with a 4-layer 3x3 BF16
 * CNN on a 9x9 grid of 4-sample displacement cells around the reference's temporal predictor,
 * refines them by exact SAD over +-2 samples for all 41 partitions, and runs the exhaustive search
 * instead where the refined 16x16 SAD exceeds fallback_ratio x the CNN's predicted SAD (a
 * non-finite fallback_ratio turns that off). References are limited to 3. Results have the layout
 * of the exhaustive search. */
#define IA_CNN_PARAMS 21988  /* f32 parameters: per layer, weights [cout][cin][3][3] then biases [cout];
                                (cin, cout) = (8, 32), (32, 32), (32, 32), (32, 4) */
#define IA_CNN_CELLS 81
#define IA_CNN_FEATURES 8
#define IA_CNN_OUTPUTS 4

/* Loads the CNN's parameters (IA_CNN_PARAMS floats, converted to BF16 on the device). */
int ia_load_proposal_weights(ia_ctx *ctx, const float *params, int n);

/* The CNN's input features [mb][IA_CNN_CELLS][IA_CNN_FEATURES] and grid centres [mb][3][2]
 * (samples) of the uploaded picture against ref_slots (1..3), for training. Synchronous. */
int ia_proposal_features(ia_ctx *ctx, const int *ref_slots, int nref, float *features, int16_t *centres);

/* The CNN's outputs [n][IA_CNN_CELLS][IA_CNN_OUTPUTS] on given features [n][IA_CNN_CELLS][IA_CNN_FEATURES]
 * (n <= macroblocks of the context), for checking. Synchronous. */
int ia_proposal_forward(ia_ctx *ctx, const float *features, int n, float *outputs);

/* The number of (macroblock, reference) pairs the last learned search sent to the exhaustive search. */
int ia_read_fallback_count(ia_ctx *ctx, int *count);

/* ---- Batching (no counterpart in the reference, whose encoder codes one stream at a time) ----
 * A batch groups contexts (streams of any sizes) so that a frame step of all of them runs as one set
 * of launches: the exhaustive search of every (stream, macroblock, reference), one persistent
 * analysis kernel whose workgroups take (stream, slice) tasks from a queue (longest first), and the
 * promotion of each stream's reconstruction to a reference (half-sample planes, chroma, records).
 * With use_graph the step is captured once as a hipGraph and replayed: the per-frame parameters
 * travel in a table copied from pinned host memory by the graph's first node. Each context's
 * results are read with ia_read_analysis / ia_read_motion after ia_batch_sync. While a batch
 * exists its contexts take no other calls but uploads, reads and reference loads between steps.
 * Contexts with IA_SEARCH_LEARNED must share their CNN weights (ia_load_proposal_weights with the
 * same parameters); their proposal runs as one persistent launch over all their macroblocks and
 * their fallbacks as one list (ia_batch_fallback_count). */
typedef struct ia_batch ia_batch;

typedef struct ia_frame {
  int qp;                 /* slice QP, 0..51 */
  int kind;               /* IA_PIC_IDR, IA_PIC_P or IA_PIC_B (CABAC streams) */
  int nref;               /* P pictures: references used, 1..nrefs; B pictures: 2 */
  int ref_slots[16];      /* P pictures: slots, newest first (as plist, enc/mod.rs:559); B pictures:
                             [0] the list 0 picture, [1] the list 1 picture (as ia_analyse_b) */
  int promote_slot;       /* -1, or the slot that receives this picture's reconstruction and records
                             as a reference when the step is done (-1 for B pictures, which are not
                             references) */
  const uint8_t *y, *cb, *cr;  /* the picture's padded planes in device memory, or null: the context's
                                  uploaded source (ia_upload_source) */
  int promote_deblock;    /* non-zero: the promoted reference is loop-filtered first, as the encoder
                             keeps its references (enc/mod.rs:653-654); 0: kept as reconstructed */
  int dist_l0, dist_l1;   /* B pictures: picture-order distances list 0 -> this picture and list 0 ->
                             list 1, 0 < dist_l0 < dist_l1 (as ia_analyse_b) */
} ia_frame;

int ia_batch_create(ia_batch **out, ia_ctx *const *ctxs, int n, int use_graph);
void ia_batch_destroy(ia_batch *b);
const char *ia_batch_last_error(const ia_batch *b);
/* Enqueues one frame step: frames[i] for context i. Asynchronous. */
int ia_batch_step(ia_batch *b, const ia_frame *frames);
/* Waits for the batch's stream. */
int ia_batch_sync(ia_batch *b);
/* The number of (stream, macroblock, reference) triples the last step sent from the learned proposal
 * to the exhaustive search. Synchronous. */
int ia_batch_fallback_count(ia_batch *b, int *count);
/* The analysis workgroups resident at once on the current device (one wave each): a batch's
 * (stream, slice) tasks run that many at a time, so slices of about equal length, about this many
 * of them over all streams, keep every worker busy. 0 if the device cannot be queried. */
int ia_batch_workers(void);

/* Tile shapes of the full-search kernel, for the autotuner (no counterpart in the reference, whose
 * search is a scalar loop): the number of compiled shapes, each one's name, and the shape a context
 * uses (-1: the default). */
int ia_fullsearch_variants(void);
const char *ia_fullsearch_variant_name(int variant);
int ia_set_fullsearch_variant(ia_ctx *ctx, int variant);

/* The loop filter of ia_promote_reference (deblock 1) on a caller's picture, allocating nothing. No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops that forbid allocation). y (mb_w*16 x mb_h*16), cb, cr (mb_w*8 x mb_h*8)
 * and the picture's records `mbs` (mb_w*mb_h) are device memory; the planes are filtered in place,
 * asynchronously on `stream` (a hipStream_t, NULL for the default stream). *stalled (device memory, may
 * be NULL) receives 1 if a row's wait for the row above gave up, else 0. The row counters live in a
 * static device array of 1 + 512 words: mb_h must be at most 512 (IA_EINVAL beyond). Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU race. */
int ia_deblock_picture_noalloc(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h,
                               int *stalled, void *stream);

#ifdef __cplusplus
}
#endif
#endif
