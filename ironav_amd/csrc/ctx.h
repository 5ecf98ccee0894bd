// ctx.h — the analysis context behind the C ABI (host side): device buffers, reference slots and
// the helpers the entry points share (api.hip, batch.hip).
#pragma once

#include <string>

#include "analyse.h"
#include "proposal.h"

using ia::Geom;

struct IaSlot {
  uint8_t *hp = nullptr;      // 4 half-sample planes, sw x sh each
  uint8_t *cb = nullptr, *cr = nullptr;
  ia_mbinfo *mbs = nullptr;   // macroblock records of the reference picture
  bool loaded = false;
};

struct ia_ctx {
  ia_config cfg;
  Geom g;
  int device = 0;
  hipStream_t own = nullptr, stream = nullptr;
  uint8_t *src[3] = {nullptr, nullptr, nullptr};
  uint8_t *tmp_luma = nullptr;
  IaSlot slots[16];
  ia_fsresult *fs = nullptr;
  int fs_nref = 0, fs_slots[16] = {0};
  bool fs_valid = false;
  int fs_variant = -1;
  // analysis outputs and working state
  uint8_t *rec[3] = {nullptr, nullptr, nullptr};
  ia_mbcode *codes = nullptr;
  ia_mbstats *stats = nullptr;
  ia_mbinfo *mbs = nullptr;
  ia::MbState *state = nullptr;
  uint32_t *entropy = nullptr;
  bool an_valid = false;
  bool an_b = false;          // the last analysis was of a B picture (not a reference)
  // learned proposal
  ia::CnnWeights *wts = nullptr;
  bool wts_loaded = false;
  uint64_t wts_hash = 0;       // of the loaded weights (a batch's learned streams share one set)
  int *fb = nullptr;           // [0], [1] counts (alternating calls), [2..] entries
  int fb_gen = 0;              // the count the next learned search uses
  float *feat = nullptr, *cnn_out = nullptr;
  int16_t *centres = nullptr;
  // loop filter: [0] stall flag, [1 .. mb_h] row progress (both 0 between launches)
  int *db = nullptr;
  std::string err;
};

namespace iah {

// the message of the last failed ia_create (no context to hold it), per thread
std::string &create_error();

inline int fail(ia_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  else create_error() = msg;
  return code;
}

inline int hip_fail(ia_ctx *c, hipError_t e, const char *what) {
  return fail(c, IA_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Checks that ref_slots[0..nref) name loaded slots of c.
int check_refs(ia_ctx *c, const int *ref_slots, int nref);

// The analysis parameters of one picture of c: QP, kind, references; src: the picture's planes in
// device memory (null: the context's uploaded source). Fails on invalid arguments or missing
// exhaustive-search results (need_fs: they must be valid for these references).
int analysis_params(ia_ctx *c, int qp, int kind, const int *ref_slots, int nref, const uint8_t *const *src,
                    bool need_fs, ia::AnParams *p);

// The analysis parameters of a B picture of c between the pictures in slots l0 (list 0) and l1
// (list 1) at picture-order distances dist_l0 (list 0 -> this picture) and dist_l1 (list 0 -> list 1);
// src as for analysis_params. Fails on invalid arguments (CABAC only, 0 < dist_l0 < dist_l1).
int b_params(ia_ctx *c, int qp, int l0, int l1, int dist_l0, int dist_l1, const uint8_t *const *src, ia::AnParams *p);

// The exhaustive search's arguments for c against ref_slots (src_y: the picture's luma, null: the
// context's uploaded source).
void search_args(ia_ctx *c, const int *ref_slots, int nref, const uint8_t *src_y, ia::FsArgs *a);

}  // namespace iah

#define IA_HIP(c, call)                                          \
  do {                                                           \
    hipError_t e_ = (call);                                      \
    if (e_ != hipSuccess) return iah::hip_fail((c), e_, #call); \
  } while (0)
