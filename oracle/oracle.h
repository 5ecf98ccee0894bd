/* oracle.h — CPU restatement of nervosys/IronAV's H.264 per-macroblock analysis, for checking.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (ironav_amd/) includes, links or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, as the checker.
 *
 * Each function names the reference file:line it follows (paths relative to the IronAV checkout;
 * enc/ = crates/ironav-codec/src/h264/enc/, h264/ = crates/ironav-codec/src/h264/).
 * Samples are held as uint16_t as the reference holds them (enc/frame.rs:319-323).
 *
 * Parity pins (what each part is checked against, in tests/):
 *  - halfpel build/predict: the decoder's interpolation (h264/inter.rs:466-547) restated separately from
 *    H.264 8.4.2.2.1 here (or_luma_mc_ref), over the vector grid of enc/subpel.rs:160-193 at 8 and 10 bits.
 *  - or_luma_mc_ref, or_chroma_mc, the nine 4x4/8x8 intra modes with the 8x8 reference filter, the
 *    16x16 and chroma modes, the inverse transforms and the loop filter (or_deblock_picture_params):
 *    the oracle's decoder (decoder.c) calls them unchanged and decodes the reference decoder's 13 CABAC
 *    fixtures (libx264 streams) to FFmpeg's pictures byte for byte, every mode and partitioning occurring
 *    (tests/test_oracle_decoder.py).
 *  - fdct4/quant/idct4, quant8/idct8: the reference's round-trip tests (enc/frame.rs:2443-2479), and
 *    closed forms (the transforms are the H.264 matrices; tests check them against numpy matrix products).
 *  - decimate_score: the known answers of enc/frame.rs:2482-2492.
 *  - entropy table: its closed form (enc/cabac.rs:162-173) recomputed in numpy.
 *  - se/ue bits: Exp-Golomb lengths (enc/bitw.rs:81-89) against the codes' definition.
 *  - full search: brute force in numpy on small pictures.
 *  - loop filter: also what H.264 8.7 fixes (identities, symmetries, reach; tests/test_oracle_deblock.py).
 *  - realtime decisions (coder_t.fast): the early P_Skip, 16x16 only, half samples only and the
 *    flat-intra rule by what they must do to a picture (tests/test_oracle_realtime.py); the shorter
 *    diamond and the integer stop parity unpinned.
 *  - the reconstruction of or_analyse_picture (prediction, dequantisation, inverse transforms), its
 *    vectors against their predictors (mvd = mv - the 8.4.1.3 predictor, P_Skip vectors by 8.4.1.1) and
 *    the CABAC bin sequence of mb_bins: the analysis of IPPP chains is written as a bitstream by the
 *    product's back end and decoded by decoder.c, which recomputes every vector from its mvd and its
 *    own predictor; the decoded pictures equal the loop-filtered reconstruction and the decoded records
 *    the analysis', and the written size is within 2% of the rate mb_bins counts (tests/test_writer.py).
 *  - B pictures (or_analyse_picture_b: direct_mvs, b_pred, the B bins, the list 1 records): the same
 *    round trip on IBBP chains written with B slices; decoder.c's B slices are pinned by the reference's
 *    five CABAC B fixtures (spatial and temporal direct, implicit weights, a B pyramid), and it
 *    recomputes every B vector from its mvd, its own predictor and its own direct derivation.
 *  - motion_search's candidate order and diamond, best_inter_ref/sub_search, code_mb's thresholds and
 *    decide_rd's order and costs: parity unpinned (the reference commits no encoder outputs and cannot be
 *    built here; any decision the syntax allows decodes correctly). Restated line by line.
 */
#ifndef IRONAV_ORACLE_H
#define IRONAV_ORACLE_H

#include <stdint.h>
#include "../include/ironav_h264a.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PAD 48

/* Half-sample planes (enc/subpel.rs:9-19): G, B, H, J, each (pw+96) x (ph+96). */
void or_halfpel_build(const uint16_t *luma, int pw, int ph, int bd, uint16_t *planes);
void or_halfpel_predict(const uint16_t *planes, int pw, int ph, int x, int y, int mvx, int mvy, int w, int h,
                        uint16_t *out);
/* The decoder's luma interpolation (h264/inter.rs:466-547), restated from H.264 8.4.2.2.1 per sample. */
void or_luma_mc_ref(const uint16_t *refp, int stride, int pw, int ph, int x, int y, int mx, int my, int w, int h,
                    uint16_t *out, int maxv);
/* Chroma eighth-sample bilinear prediction (h264/inter.rs:404-422). */
void or_chroma_mc(const uint16_t *refp, int pw, int ph, int x, int y, int mx, int my, int w, int h, uint16_t *out);

/* Primitives (enc/frame.rs, h264/idct.rs), exposed for the known-answer tests. */
void or_fdct4(const int32_t *d, int32_t *out);
void or_fdct8(const int32_t *d, int32_t *out);
void or_idct4_add(const int32_t *c, uint16_t *dst, int stride, int maxv);
void or_idct8_add(const int32_t *c, uint16_t *dst, int stride, int maxv);
void or_hadamard4(const int32_t *c, int32_t *out);
int32_t or_quant(int32_t c, int32_t mf, int32_t f, int qbits);
void or_quant8(const int32_t *src, const int32_t *pred, int qp, int intra, int32_t *lv, int32_t *co);
int32_t or_satd4(const int32_t *d);
int32_t or_decimate_score(const int32_t *lv);
int32_t or_decimate_score8(const int32_t *lv);
uint32_t or_ue_bits(uint32_t v);
uint32_t or_se_bits(int32_t v);
int32_t or_lambda(int qp);
void or_entropy_table(uint32_t *out /* 64*2 */);
void or_predict_nxn(int n, int mode, const int32_t *top16, const int32_t *left8, int32_t topleft, int have_top,
                    int have_left, int have_topleft, uint16_t *out, int bd);
void or_filter_8x8_edges(int32_t *top16, int32_t *left8, int32_t *topleft, int have_top, int have_left,
                         int have_topleft);
void or_predict_16x16(int mode, const int32_t *top, const int32_t *left, int32_t topleft, int have_top,
                      int have_left, uint16_t *out, int bd);
void or_predict_chroma(int mode, const int32_t *top8, const int32_t *left16, int32_t topleft, int have_top,
                       int have_left, uint16_t *out, int bd, int h);

/* Exhaustive integer SAD search, brute force, with the window and order of ia_motion_search.
 * src_y: padded luma (stride mb_w*16); ref_y[r]: deblocked reference luma (same layout).
 * out: [nref][mb][IA_NPART]. */
void or_fullsearch(const uint8_t *src_y, const uint8_t *const *ref_y, int nref, int mb_w, int mb_h, int range,
                   ia_fsresult *out);

/* The loop filter the encoder applies to each reconstruction before it becomes a reference
 * (enc/mod.rs:653-654, h264/deblock.rs:410-421 and :608-695): frame macroblocks, 4:2:0, 8-bit,
 * filtering across slice edges with zero offsets. In place on padded planes y (stride mb_w*16),
 * cb and cr (stride mb_w*8), from the picture's macroblock records. */
void or_deblock_picture(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h);
/* The same with the settings of each slice (indexed by ia_mbinfo.slice - 1): disable_deblocking_filter_idc,
 * FilterOffsetA/B (slice_alpha/beta_c0_offset_div2 * 2) and the chroma QP offsets of Cb and Cr. */
typedef struct or_slice_filter {
  int disable_idc, alpha_offset, beta_offset, chroma_qp_offset[2];
} or_slice_filter;
void or_deblock_picture_params(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h,
                               const or_slice_filter *sp);

/* sizeof of ia_config, ia_mbinfo, ia_mbcode, ia_mbstats and ia_fsresult, for the binding tests. */
void or_abi_sizes(int32_t *out);

/* One picture's analysis across all slices: the oracle of ia_analyse_picture. Inputs are 8-bit
 * padded planes; references are given by their deblocked planes and macroblock records (mbs may be
 * NULL). fs: full-search winners [nref][mb][IA_NPART] for IA_SEARCH_FULL, else NULL. Returns 0, or
 * -1 on an invalid argument. */
int or_analyse_picture(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                       int kind, const uint8_t *const *ref_y, const uint8_t *const *ref_cb,
                       const uint8_t *const *ref_cr, const ia_mbinfo *const *ref_mbs, int nref,
                       const ia_fsresult *fs, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y,
                       uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs);

/* The same with the slices coded on `threads` host threads (slice s on thread s mod threads), as the
 * reference's encoder codes each slice on a thread of its own (enc/mod.rs:623-627). Returns -2 if a
 * thread could not be started. */
int or_analyse_picture_mt(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                          int kind, const uint8_t *const *ref_y, const uint8_t *const *ref_cb,
                          const uint8_t *const *ref_cr, const ia_mbinfo *const *ref_mbs, int nref,
                          const ia_fsresult *fs, ia_mbcode *codes, ia_mbstats *stats, uint8_t *recon_y,
                          uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs, int threads);

/* One B picture's analysis (enc/frame.rs:1518-1780 per macroblock, the slices of code_picture with
 * `with_b`, enc/mod.rs:551-566,619-621): ref_*[0] is the list 0 picture (the nearest reference shown
 * before it), ref_*[1] the list 1 picture (the nearest shown after it; its records, ref_mbs[1], give
 * the co-located motion); tb and td its picture-order distances from the list 0 picture to this one
 * and to the list 1 picture. CABAC only. Returns 0, or -1 on an invalid argument. */
int or_analyse_picture_b(const ia_config *cfg, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp,
                         const uint8_t *const *ref_y, const uint8_t *const *ref_cb, const uint8_t *const *ref_cr,
                         const ia_mbinfo *const *ref_mbs, int tb, int td, ia_mbcode *codes, ia_mbstats *stats,
                         uint8_t *recon_y, uint8_t *recon_cb, uint8_t *recon_cr, ia_mbinfo *mbs, int threads);

#ifdef __cplusplus
}
#endif
#endif
