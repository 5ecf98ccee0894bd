// common.h — shared definitions of the gfx950 kernels and the host side of the H.264 analysis library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ironav_h264a.h"

namespace ia {

// Border of the half-sample planes: a block 24 samples outside the picture plus the 6-tap reach
// (crates/ironav-codec/src/h264/enc/subpel.rs:7).
constexpr int PAD = 48;

// Geometry of one picture size.
struct Geom {
  int mb_w, mb_h;  // macroblocks
  int pw, ph;      // padded luma size
  int sw, sh;      // half-sample plane size (pw + 2*PAD, ph + 2*PAD)
  int cw, ch;      // chroma plane size
};

inline Geom make_geom(int width, int height) {
  Geom g;
  g.mb_w = (width + 15) / 16;
  g.mb_h = (height + 15) / 16;
  g.pw = g.mb_w * 16;
  g.ph = g.mb_h * 16;
  g.sw = g.pw + 2 * PAD;
  g.sh = g.ph + 2 * PAD;
  g.cw = g.pw / 2;
  g.ch = g.ph / 2;
  return g;
}

// Launchers, one per kernel file. Each returns the hipError_t of its launch.
hipError_t launch_halfpel(const uint8_t *luma, int pw, int ph, uint8_t *planes, hipStream_t stream);

struct FsArgs {
  const uint8_t *src_y;       // padded luma of the picture, stride pw
  const uint8_t *ref_g[16];   // integer plane G of each reference (stride sw, PAD border)
  ia_fsresult *out;           // [nref][mb][IA_NPART]
  int nref, mb_w, mb_h, range;
};
// `variant` selects a tile shape of the search (see fullsearch.hip); -1 picks the default.
hipError_t launch_fullsearch(const FsArgs &a, int variant, hipStream_t stream);
int fullsearch_variants();
const char *fullsearch_variant_name(int v);
int fullsearch_default_variant();
// The search of many pictures in one launch (args in device memory): member k of the launch
// (offs[k] <= block < offs[k+1], each member's block count 8 * ceil(nmb / 8) * its most references)
// searches picture sidx[k]; all members' ranges are at most max_range and share the tile `variant`.
hipError_t launch_fullsearch_batch(const FsArgs *args, const int *offs, const int *sidx, int nmem, int nblocks,
                                   int variant, int max_range, hipStream_t stream);

// One picture's half-sample build in a batched launch: tiles [first, first + tiles) of the launch
// build `planes` from `luma`; a null `planes` skips the picture this time.
struct HpJob {
  const uint8_t *luma;
  uint8_t *planes;
  int pw, ph, tiles_x, first;
};
// Strip of the half-sample kernels: one wave, 256 columns x HP_TH rows. A single picture's launch
// has at most about HP_WAVES waves; a wave builds one or more strips of one column band.
constexpr int HP_TW = 256, HP_TH = 4, HP_WAVES = 8192;
hipError_t launch_halfpel_batch(const HpJob *jobs, int njobs, int nblocks, hipStream_t stream);

// The loop filter of one picture in place (one wave per macroblock row): planes y (stride mb_w*16),
// cb and cr (stride mb_w*8) from the picture's records. prog: mb_h counters, 0 before the launch and
// 0 again after it; err: set non-zero if a row stalled waiting for the one above.
hipError_t launch_deblock(uint8_t *y, uint8_t *cb, uint8_t *cr, const ia_mbinfo *mbs, int mb_w, int mb_h, int *prog,
                          int *err, hipStream_t stream);
// One picture's loop filter in a batched launch: blocks [first, first + mb_h) filter its rows; a null
// `y` skips the picture this time.
struct DbJob {
  uint8_t *y, *cb, *cr;
  const ia_mbinfo *mbs;
  int *prog;
  int mb_w, mb_h, first;
};
hipError_t launch_deblock_batch(const DbJob *jobs, int njobs, int nblocks, int *err, hipStream_t stream);

}  // namespace ia
