// analyse.h — parameters of the per-macroblock analysis kernel (analyse.hip) and its state records.
#pragma once

#include "common.h"
#include "h264_tables.h"

namespace ia {

// What later macroblocks of a slice need to know about a coded one (enc/frame.rs:239-272).
struct MbState {
  uint8_t i4, inter, t8x8, skip, kind, cbp, chroma_mode, cbf_dc;
  uint8_t modes[16], nnz[16], cnnz[2][4];
  int16_t mvs[16][2];
  int8_t refs[4];
  uint8_t mvd[16][2];
  // list 1 of a B macroblock (vectors, reference index per 8x8 block, -1 unused; |mvd|), and whether
  // it is B_Skip or B_Direct_16x16
  int16_t mvs1[16][2];
  int8_t refs1[4];
  uint8_t mvd1[16][2];
  uint8_t bdirect, pad[3];
};

struct AnParams {
  const uint8_t *src[3];        // padded source planes
  const uint8_t *hp[16];        // half-sample planes of each list-0 reference
  const uint8_t *rcb[16], *rcr[16];
  const ia_mbinfo *rmbs0;       // records of the co-located picture (P: reference 0; B: the list 1 picture), may be null
  const ia_fsresult *fs;        // exhaustive-search winners [nref][mb][IA_NPART], or null
  uint8_t *rec[3];              // reconstruction before deblocking
  ia_mbcode *codes;
  ia_mbstats *stats;
  ia_mbinfo *mbs;
  MbState *state;               // per macroblock, for the neighbours of later ones
  const uint32_t *entropy;      // [64][2] CABAC bit costs in 1/256 bits
  int mb_w, mb_h, nslices, nref;
  int qp_slice;                 // slice QP
  int qp;                       // QP'Y of the transform (qp_slice at 8 bits)
  int qpc;                      // QP'C
  int lambda;
  int p_slice, sub8, use_cabac, transform8;
  int b_slice;                  // B picture: hp/rcb/rcr [0] the list 0 picture, [1] the list 1 picture
  int dist_tb, dist_td;         // B: picture-order distances list 0 -> this picture, list 0 -> list 1
  double lam2;                  // 1.6 * 2^((QP'Y - 12) / 3)
  double skip_thresh;           // 256 * 0.625 * 2^(QP'Y / 6)
  int fast;                     // realtime decisions (enc/frame.rs:348)
  double fast_skip_y;           // 256 * 0.05 * 2^(QP'Y / 6): luma SAD bound of the untransformed P_Skip
  double fast_skip_c;           // 128 * 0.05 * 2^(QP'Y / 6): chroma SAD bound of the same
  double fast_flat;             // 256 * 0.8 * 2^(QP'Y / 6): Intra16x16 cost below which Intra4x4 is not tried
};

hipError_t launch_analyse(const AnParams &p, hipStream_t stream);
// Slices of many pictures: tasks[t] = picture << 16 | slice (t < ntasks) with ps[picture] in device
// memory; `workers` workgroups take tasks from a queue counted in *head (0 at the launch).
hipError_t launch_analyse_batch(const AnParams *ps, const int *tasks, int ntasks, int *head, int workers,
                                hipStream_t stream);
// workgroups of the batched analysis resident on the device at once (0 on error)
int analyse_batch_workers();

}  // namespace ia
