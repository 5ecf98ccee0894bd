// proposal.h — the learned motion proposal: a small CNN on MFMA proposing candidate vectors that an
// exact SAD refinement turns into per-partition winners (the format of the exhaustive search).
#pragma once

#include "common.h"

namespace ia {

// The CNN runs on a 9x9 grid of displacement cells per macroblock: cell (cx, cy) is the displacement
// centre + 4 * (cx - 4, cy - 4) samples, centre = the reference's temporal predictor.
constexpr int CNN_GRID = 9, CNN_CELLS = 81, CNN_CH = 32, CNN_IN = 8, CNN_OUT = 4, CNN_TOPK = 8;
constexpr int CNN_LAYERS = 4;

// Weights in the layout the MFMA operands read: bf16 bits, [layer][tap][cout][cin] (cin contiguous);
// layer 0 reads CNN_IN real input channels (the rest zero), layer 3 writes CNN_OUT real outputs.
struct CnnWeights {
  uint16_t w[CNN_LAYERS][9][CNN_CH][CNN_CH];
  float b[CNN_LAYERS][CNN_CH];
};

struct PropArgs {
  const uint8_t *src_y;         // padded luma of the picture
  const uint8_t *ref_g[3];      // integer planes G of references 0..nref-1 (PAD border)
  const ia_mbinfo *rmbs0;       // reference 0's macroblock records (temporal predictors), may be null
  const CnnWeights *wts;
  ia_fsresult *out;             // [nref][mb][IA_NPART]
  int *fb_list;                 // fallback (fb_tag | mb * 4 + ref) entries; null: no fallback
  int fb_tag;                   // added to the entries (the picture's index in a batch, << 20)
  int *fb_count;
  float *feat;                  // optional: CNN input features [mb][81][CNN_IN] (for training)
  float *cnn_out;               // optional: CNN outputs [mb][81][CNN_OUT]
  int16_t *centre;              // optional: window centre per [mb][3] (x, y) in samples
  const float *feat_in;         // optional: features to run the CNN on instead of computing them
  int mb_w, mb_h, nref;
  float fallback_ratio;
  int features_only;            // 1: compute features (and centres) only
};

hipError_t launch_proposal(const PropArgs &a, hipStream_t stream);
// Many pictures in one launch (args in device memory): macroblock g < total of the launch is
// macroblock g - offs[k] of picture sidx[k] (offs[k] <= g < offs[k+1]); all share args[sidx[0]].wts.
hipError_t launch_proposal_batch(const PropArgs *args, const int *offs, const int *sidx, int nmem, int total,
                                 hipStream_t stream);
// the exhaustive search over the (mb, ref) entries of a fallback list; the search also sets *reset
// to 0 (the count the next list will be gathered in)
hipError_t launch_fullsearch_list(const FsArgs &a, const int *list, const int *count, int *reset, hipStream_t stream);
// the fallback list of a batch: entries picture << 20 | mb << 2 | ref with args[picture] in device
// memory (ranges at most max_range; at most max_entries entries)
hipError_t launch_fullsearch_list_batch(const FsArgs *args, const int *list, const int *count, int max_range,
                                        int max_entries, hipStream_t stream);

}  // namespace ia
