// api.hip — the C ABI of include/ironav_h264a.h: context, device buffers and launches.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>

#include "ctx.h"

namespace iah {

std::string &create_error() {
  static thread_local std::string e;
  return e;
}

int check_refs(ia_ctx *c, const int *ref_slots, int nref) {
  if (nref < 1 || nref > c->cfg.nrefs || !ref_slots) return fail(c, IA_EINVAL, "nref must be within 1..nrefs");
  for (int r = 0; r < nref; r++) {
    if (ref_slots[r] < 0 || ref_slots[r] >= c->cfg.nrefs) return fail(c, IA_EINVAL, "reference slot out of range");
    if (!c->slots[ref_slots[r]].loaded) return fail(c, IA_ESTATE, "reference slot is empty");
  }
  return IA_OK;
}

void search_args(ia_ctx *c, const int *ref_slots, int nref, const uint8_t *src_y, ia::FsArgs *a) {
  memset(a, 0, sizeof *a);
  a->src_y = src_y ? src_y : c->src[0];
  for (int r = 0; r < nref; r++) a->ref_g[r] = c->slots[ref_slots[r]].hp;
  a->out = c->fs;
  a->nref = nref;
  a->mb_w = c->g.mb_w;
  a->mb_h = c->g.mb_h;
  a->range = c->cfg.search_range;
}

static int chroma_qp(int qp) {
  static const uint8_t T[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
  const int q = qp < 0 ? 0 : (qp > 51 ? 51 : qp);
  return q < 30 ? q : T[q - 30];
}

int analysis_params(ia_ctx *c, int qp, int kind, const int *ref_slots, int nref, const uint8_t *const *src,
                    bool need_fs, ia::AnParams *pp) {
  ia::AnParams &p = *pp;
  if (qp < 0 || qp > 51) return fail(c, IA_EINVAL, "qp must be within 0..51");
  if (kind != IA_PIC_IDR && kind != IA_PIC_P) return fail(c, IA_EINVAL, "kind must be IA_PIC_IDR or IA_PIC_P");
  memset(&p, 0, sizeof p);
  const bool pslice = kind == IA_PIC_P;
  if (pslice) {
    int rc = check_refs(c, ref_slots, nref);
    if (rc) return rc;
  } else {
    nref = 0;
  }
  if (pslice && c->cfg.search_mode != IA_SEARCH_DIAMOND) {
    if (need_fs) {
      if (!c->fs_valid || c->fs_nref != nref) return fail(c, IA_ESTATE, "no exhaustive-search results for these references");
      for (int r = 0; r < nref; r++)
        if (c->fs_slots[r] != ref_slots[r]) return fail(c, IA_ESTATE, "exhaustive-search results are for other references");
    }
    p.fs = c->fs;
  }
  for (int k = 0; k < 3; k++) {
    p.src[k] = src && src[k] ? src[k] : c->src[k];
    p.rec[k] = c->rec[k];
  }
  for (int r = 0; r < nref; r++) {
    const IaSlot &sl = c->slots[ref_slots[r]];
    p.hp[r] = sl.hp;
    p.rcb[r] = sl.cb;
    p.rcr[r] = sl.cr;
  }
  p.rmbs0 = nref > 0 ? c->slots[ref_slots[0]].mbs : nullptr;
  p.codes = c->codes;
  p.stats = c->stats;
  p.mbs = c->mbs;
  p.state = c->state;
  p.entropy = c->entropy;
  p.mb_w = c->g.mb_w;
  p.mb_h = c->g.mb_h;
  p.nslices = c->cfg.slices < c->g.mb_h ? c->cfg.slices : c->g.mb_h;
  p.nref = nref;
  p.qp_slice = qp;
  p.qp = qp;  // + 6 * (bit_depth - 8) at 8 bits
  p.qpc = chroma_qp(qp);
  // lambda = max(1, round(2^((qp - 12) / 6))), enc/frame.rs:164-166
  p.lambda = (int)round(pow(2.0, ((double)p.qp - 12.0) / 6.0));
  if (p.lambda < 1) p.lambda = 1;
  p.p_slice = pslice;
  p.sub8 = c->cfg.sub8x8 && qp <= 20;
  p.use_cabac = c->cfg.cabac != 0;
  p.transform8 = c->cfg.cabac && c->cfg.transform8;
  p.lam2 = 1.6 * pow(2.0, ((double)p.qp - 12.0) / 3.0);
  p.skip_thresh = 256.0 * 0.625 * pow(2.0, (double)p.qp / 6.0);
  const double step = pow(2.0, (double)p.qp / 6.0);
  p.fast = c->cfg.fast != 0;
  p.fast_skip_y = 256.0 * 0.05 * step;
  p.fast_skip_c = 128.0 * 0.05 * step;
  p.fast_flat = 256.0 * 0.8 * step;
  return IA_OK;
}

int b_params(ia_ctx *c, int qp, int l0, int l1, int dist_l0, int dist_l1, const uint8_t *const *src, ia::AnParams *pp) {
  if (!c->cfg.cabac) return fail(c, IA_EUNSUPPORTED, "B pictures are coded with CABAC only");
  if (dist_l0 <= 0 || dist_l1 <= dist_l0)
    return fail(c, IA_EINVAL, "the B picture must lie between its references: 0 < dist_l0 < dist_l1");
  const int slots[2] = {l0, l1};
  int rc = analysis_params(c, qp, IA_PIC_P, slots, 2, src, false, pp);
  if (rc) return rc;
  // one picture per list: planes [0] list 0, [1] list 1; the co-located records are list 1's
  ia::AnParams &p = *pp;
  p.fs = nullptr;
  p.nref = 1;
  p.b_slice = 1;
  p.dist_tb = dist_l0;
  p.dist_td = dist_l1;
  p.rmbs0 = c->slots[l1].mbs;
  return IA_OK;
}

}  // namespace iah

using iah::fail;
using iah::hip_fail;

namespace {

int check_config(const ia_config *cfg, std::string &why) {
  if (cfg->width < 16 || cfg->height < 16 || cfg->width > 8192 || cfg->height > 8192 || (cfg->width & 1) || (cfg->height & 1)) {
    why = "width and height must be even and within 16..8192";
    return IA_EINVAL;
  }
  if (cfg->bit_depth != 8) {
    why = "only 8-bit pictures are supported on the GPU";
    return IA_EUNSUPPORTED;
  }
  if (cfg->nrefs < 1 || cfg->nrefs > 16) {
    why = "nrefs must be within 1..16";
    return IA_EINVAL;
  }
  int mb_h = (cfg->height + 15) / 16;
  if (cfg->slices < 1 || cfg->slices > mb_h) {
    why = "slices must be within 1..mb_h";
    return IA_EINVAL;
  }
  if (cfg->search_mode < IA_SEARCH_DIAMOND || cfg->search_mode > IA_SEARCH_LEARNED) {
    why = "unknown search_mode";
    return IA_EINVAL;
  }
  if (cfg->search_range < 1 || cfg->search_range > 64) {
    why = "search_range must be within 1..64";
    return IA_EINVAL;
  }
  return IA_OK;
}

}  // namespace

extern "C" {

int ia_create(ia_ctx **out, const ia_config *cfg) {
  if (!out || !cfg) return fail(nullptr, IA_EINVAL, "null argument");
  *out = nullptr;
  std::string why;
  int rc = check_config(cfg, why);
  if (rc != IA_OK) return fail(nullptr, rc, why);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, IA_EDEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return fail(nullptr, IA_EDEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, IA_EDEVICE, std::string("the library is built for gfx950; device is ") + prop.gcnArchName);
  ia_ctx *c = new ia_ctx();
  c->cfg = *cfg;
  c->g = ia::make_geom(cfg->width, cfg->height);
  c->device = dev;
  const Geom &g = c->g;
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch, hplen = (size_t)4 * g.sw * g.sh;
  const size_t nmb = (size_t)g.mb_w * g.mb_h;
  auto alloc = [&](void **p, size_t n) { return hipMalloc(p, n); };
  bool ok = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) == hipSuccess;
  c->stream = c->own;
  ok = ok && alloc((void **)&c->src[0], ylen) == hipSuccess && alloc((void **)&c->src[1], clen) == hipSuccess &&
       alloc((void **)&c->src[2], clen) == hipSuccess && alloc((void **)&c->tmp_luma, ylen) == hipSuccess &&
       alloc((void **)&c->fs, (size_t)cfg->nrefs * nmb * IA_NPART * sizeof(ia_fsresult)) == hipSuccess;
  ok = ok && alloc((void **)&c->rec[0], ylen) == hipSuccess && alloc((void **)&c->rec[1], clen) == hipSuccess &&
       alloc((void **)&c->rec[2], clen) == hipSuccess && alloc((void **)&c->codes, nmb * sizeof(ia_mbcode)) == hipSuccess &&
       alloc((void **)&c->stats, nmb * sizeof(ia_mbstats)) == hipSuccess &&
       alloc((void **)&c->mbs, nmb * sizeof(ia_mbinfo)) == hipSuccess &&
       alloc((void **)&c->state, nmb * sizeof(ia::MbState)) == hipSuccess &&
       alloc((void **)&c->entropy, 128 * sizeof(uint32_t)) == hipSuccess &&
       alloc((void **)&c->wts, sizeof(ia::CnnWeights)) == hipSuccess &&
       alloc((void **)&c->fb, (2 + 4 * nmb) * sizeof(int)) == hipSuccess &&
       alloc((void **)&c->db, (1 + (size_t)g.mb_h) * sizeof(int)) == hipSuccess;
  ok = ok && hipMemset(c->fb, 0, 2 * sizeof(int)) == hipSuccess &&
       hipMemset(c->db, 0, (1 + (size_t)g.mb_h) * sizeof(int)) == hipSuccess;
  if (ok) {
    // -log2 of the MPS/LPS probability per CABAC state in 1/256 bits (enc/cabac.rs:162-173), computed
    // with the C library's pow/log2 as the reference computes it with Rust's (which call the same)
    uint32_t ent[128];
    for (int p = 0; p < 64; p++) {
      const double lps = 0.5 * pow(0.01875 / 0.5, (double)p / 63.0);
      ent[2 * p] = (uint32_t)round(-log2(1.0 - lps) * 256.0);
      ent[2 * p + 1] = (uint32_t)round(-log2(lps) * 256.0);
    }
    ok = hipMemcpy(c->entropy, ent, sizeof ent, hipMemcpyHostToDevice) == hipSuccess;
  }
  for (int s = 0; ok && s < cfg->nrefs; s++) {
    IaSlot &sl = c->slots[s];
    ok = alloc((void **)&sl.hp, hplen) == hipSuccess && alloc((void **)&sl.cb, clen) == hipSuccess &&
         alloc((void **)&sl.cr, clen) == hipSuccess && alloc((void **)&sl.mbs, nmb * sizeof(ia_mbinfo)) == hipSuccess;
  }
  if (!ok) {
    ia_destroy(c);
    return fail(nullptr, IA_ENOMEM, "device allocation failed");
  }
  *out = c;
  return IA_OK;
}

void ia_destroy(ia_ctx *c) {
  if (!c) return;
  if (c->own) (void)hipStreamSynchronize(c->own);
  if (c->stream && c->stream != c->own) (void)hipStreamSynchronize(c->stream);
  for (auto *p : c->src) (void)hipFree(p);
  for (auto *p : c->rec) (void)hipFree(p);
  (void)hipFree(c->codes);
  (void)hipFree(c->stats);
  (void)hipFree(c->mbs);
  (void)hipFree(c->state);
  (void)hipFree(c->entropy);
  (void)hipFree(c->wts);
  (void)hipFree(c->fb);
  (void)hipFree(c->db);
  (void)hipFree(c->feat);
  (void)hipFree(c->cnn_out);
  (void)hipFree(c->centres);
  (void)hipFree(c->tmp_luma);
  (void)hipFree(c->fs);
  for (auto &s : c->slots) {
    (void)hipFree(s.hp);
    (void)hipFree(s.cb);
    (void)hipFree(s.cr);
    (void)hipFree(s.mbs);
  }
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

const char *ia_last_error(const ia_ctx *c) { return c ? c->err.c_str() : iah::create_error().c_str(); }

int ia_set_stream(ia_ctx *c, void *hip_stream) {
  if (!c) return IA_EINVAL;
  c->stream = hip_stream ? (hipStream_t)hip_stream : c->own;
  return IA_OK;
}

int ia_set_reference(ia_ctx *c, int slot, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, const ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  if (slot < 0 || slot >= c->cfg.nrefs || !y || !cb || !cr) return fail(c, IA_EINVAL, "bad reference slot or plane");
  const Geom &g = c->g;
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch, nmb = (size_t)g.mb_w * g.mb_h;
  IaSlot &s = c->slots[slot];
  IA_HIP(c, hipMemcpyAsync(c->tmp_luma, y, ylen, hipMemcpyHostToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(s.cb, cb, clen, hipMemcpyHostToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(s.cr, cr, clen, hipMemcpyHostToDevice, c->stream));
  if (mbs) IA_HIP(c, hipMemcpyAsync(s.mbs, mbs, nmb * sizeof(ia_mbinfo), hipMemcpyHostToDevice, c->stream));
  else IA_HIP(c, hipMemsetAsync(s.mbs, 0, nmb * sizeof(ia_mbinfo), c->stream));
  IA_HIP(c, ia::launch_halfpel(c->tmp_luma, g.pw, g.ph, s.hp, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  s.loaded = true;
  c->fs_valid = false;
  return IA_OK;
}

int ia_promote_reference(ia_ctx *c, int slot, int deblock) {
  if (!c) return IA_EINVAL;
  if (slot < 0 || slot >= c->cfg.nrefs) return fail(c, IA_EINVAL, "bad reference slot");
  if (!c->an_valid) return fail(c, IA_ESTATE, "no analysed picture to promote");
  // B pictures are never kept as references (enc/mod.rs:656-667)
  if (c->an_b) return fail(c, IA_ESTATE, "a B picture is not a reference");
  const Geom &g = c->g;
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch, nmb = (size_t)g.mb_w * g.mb_h;
  IaSlot &s = c->slots[slot];
  s.loaded = false;
  IA_HIP(c, hipMemcpyAsync(c->tmp_luma, c->rec[0], ylen, hipMemcpyDeviceToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(s.cb, c->rec[1], clen, hipMemcpyDeviceToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(s.cr, c->rec[2], clen, hipMemcpyDeviceToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(s.mbs, c->mbs, nmb * sizeof(ia_mbinfo), hipMemcpyDeviceToDevice, c->stream));
  if (deblock) IA_HIP(c, ia::launch_deblock(c->tmp_luma, s.cb, s.cr, c->mbs, g.mb_w, g.mb_h, c->db + 1, c->db, c->stream));
  IA_HIP(c, ia::launch_halfpel(c->tmp_luma, g.pw, g.ph, s.hp, c->stream));
  int stalled = 0;
  IA_HIP(c, hipMemcpyAsync(&stalled, c->db, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  if (stalled) {
    IA_HIP(c, hipMemset(c->db, 0, (1 + (size_t)g.mb_h) * sizeof(int)));
    return fail(c, IA_EDEVICE, "the loop filter's rows stalled");
  }
  s.loaded = true;
  c->fs_valid = false;
  return IA_OK;
}

int ia_read_reference(ia_ctx *c, int slot, uint8_t *y, uint8_t *cb, uint8_t *cr, ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  if (slot < 0 || slot >= c->cfg.nrefs) return fail(c, IA_EINVAL, "bad reference slot");
  if (!c->slots[slot].loaded) return fail(c, IA_ESTATE, "reference slot is empty");
  const Geom &g = c->g;
  const IaSlot &s = c->slots[slot];
  const size_t clen = (size_t)g.cw * g.ch, nmb = (size_t)g.mb_w * g.mb_h;
  // the luma is plane G's inside (its border repeats the picture's edges)
  if (y)
    IA_HIP(c, hipMemcpy2DAsync(y, g.pw, s.hp + (size_t)ia::PAD * g.sw + ia::PAD, g.sw, g.pw, g.ph, hipMemcpyDeviceToHost,
                               c->stream));
  if (cb) IA_HIP(c, hipMemcpyAsync(cb, s.cb, clen, hipMemcpyDeviceToHost, c->stream));
  if (cr) IA_HIP(c, hipMemcpyAsync(cr, s.cr, clen, hipMemcpyDeviceToHost, c->stream));
  if (mbs) IA_HIP(c, hipMemcpyAsync(mbs, s.mbs, nmb * sizeof(ia_mbinfo), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_read_halfpel(ia_ctx *c, int slot, uint8_t *out) {
  if (!c) return IA_EINVAL;
  if (slot < 0 || slot >= c->cfg.nrefs || !out) return fail(c, IA_EINVAL, "bad slot or output");
  if (!c->slots[slot].loaded) return fail(c, IA_ESTATE, "reference slot is empty");
  IA_HIP(c, hipMemcpyAsync(out, c->slots[slot].hp, (size_t)4 * c->g.sw * c->g.sh, hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_upload_source(ia_ctx *c, const uint8_t *y, const uint8_t *cb, const uint8_t *cr) {
  if (!c) return IA_EINVAL;
  if (!y || !cb || !cr) return fail(c, IA_EINVAL, "null plane");
  const Geom &g = c->g;
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch;
  IA_HIP(c, hipMemcpyAsync(c->src[0], y, ylen, hipMemcpyHostToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(c->src[1], cb, clen, hipMemcpyHostToDevice, c->stream));
  IA_HIP(c, hipMemcpyAsync(c->src[2], cr, clen, hipMemcpyHostToDevice, c->stream));
  c->fs_valid = false;
  return IA_OK;
}

int ia_motion_search(ia_ctx *c, const int *ref_slots, int nref) {
  if (!c) return IA_EINVAL;
  int rc = iah::check_refs(c, ref_slots, nref);
  if (rc) return rc;
  if (c->cfg.search_mode == IA_SEARCH_LEARNED) {
    if (!c->wts_loaded) return fail(c, IA_ESTATE, "the proposal's weights are not loaded");
    if (nref > 3) return fail(c, IA_EUNSUPPORTED, "the learned proposal takes at most 3 references");
    ia::PropArgs p;
    memset(&p, 0, sizeof p);
    p.src_y = c->src[0];
    for (int r = 0; r < nref; r++) p.ref_g[r] = c->slots[ref_slots[r]].hp;
    p.rmbs0 = c->slots[ref_slots[0]].mbs;
    p.wts = c->wts;
    p.out = c->fs;
    // the fallback count alternates between two words: each call's list search clears the other
    // one, which the next call counts in (a non-finite ratio turns the fallback off)
    const int gen = c->fb_gen;
    const bool fallback = std::isfinite(c->cfg.fallback_ratio);
    p.fb_count = c->fb + gen;
    p.fb_list = fallback ? c->fb + 2 : nullptr;
    p.mb_w = c->g.mb_w;
    p.mb_h = c->g.mb_h;
    p.nref = nref;
    p.fallback_ratio = c->cfg.fallback_ratio;
    IA_HIP(c, ia::launch_proposal(p, c->stream));
    ia::FsArgs a;
    iah::search_args(c, ref_slots, nref, nullptr, &a);
    if (fallback) {
      IA_HIP(c, ia::launch_fullsearch_list(a, c->fb + 2, c->fb + gen, c->fb + (gen ^ 1), c->stream));
      c->fb_gen = gen ^ 1;
    }
    c->fs_nref = nref;
    for (int r = 0; r < nref; r++) c->fs_slots[r] = ref_slots[r];
    c->fs_valid = true;
    return IA_OK;
  }
  ia::FsArgs a;
  iah::search_args(c, ref_slots, nref, nullptr, &a);
  IA_HIP(c, ia::launch_fullsearch(a, c->fs_variant, c->stream));
  c->fs_nref = nref;
  for (int r = 0; r < nref; r++) c->fs_slots[r] = ref_slots[r];
  c->fs_valid = true;
  return IA_OK;
}

int ia_read_motion(ia_ctx *c, ia_fsresult *out) {
  if (!c) return IA_EINVAL;
  if (!out) return fail(c, IA_EINVAL, "null output");
  if (!c->fs_valid) return fail(c, IA_ESTATE, "no motion search results");
  const size_t n = (size_t)c->fs_nref * c->g.mb_w * c->g.mb_h * IA_NPART;
  IA_HIP(c, hipMemcpyAsync(out, c->fs, n * sizeof(ia_fsresult), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}


int ia_analyse_uploaded(ia_ctx *c, int qp, int kind, const int *ref_slots, int nref) {
  if (!c) return IA_EINVAL;
  ia::AnParams p;
  int rc = iah::analysis_params(c, qp, kind, ref_slots, nref, nullptr, true, &p);
  if (rc) return rc;
  IA_HIP(c, ia::launch_analyse(p, c->stream));
  c->an_valid = true;
  c->an_b = false;
  return IA_OK;
}

int ia_analyse_b(ia_ctx *c, int qp, int l0_slot, int l1_slot, int dist_l0, int dist_l1) {
  if (!c) return IA_EINVAL;
  ia::AnParams p;
  int rc = iah::b_params(c, qp, l0_slot, l1_slot, dist_l0, dist_l1, nullptr, &p);
  if (rc) return rc;
  IA_HIP(c, ia::launch_analyse(p, c->stream));
  c->an_valid = true;
  c->an_b = true;
  return IA_OK;
}

int ia_read_analysis(ia_ctx *c, ia_mbcode *codes, ia_mbstats *stats, uint8_t *ry, uint8_t *rcb, uint8_t *rcr,
                     ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  if (!c->an_valid) return fail(c, IA_ESTATE, "no analysis results");
  const Geom &g = c->g;
  const size_t ylen = (size_t)g.pw * g.ph, clen = (size_t)g.cw * g.ch, nmb = (size_t)g.mb_w * g.mb_h;
  if (codes) IA_HIP(c, hipMemcpyAsync(codes, c->codes, nmb * sizeof(ia_mbcode), hipMemcpyDeviceToHost, c->stream));
  if (stats) IA_HIP(c, hipMemcpyAsync(stats, c->stats, nmb * sizeof(ia_mbstats), hipMemcpyDeviceToHost, c->stream));
  if (mbs) IA_HIP(c, hipMemcpyAsync(mbs, c->mbs, nmb * sizeof(ia_mbinfo), hipMemcpyDeviceToHost, c->stream));
  if (ry) IA_HIP(c, hipMemcpyAsync(ry, c->rec[0], ylen, hipMemcpyDeviceToHost, c->stream));
  if (rcb) IA_HIP(c, hipMemcpyAsync(rcb, c->rec[1], clen, hipMemcpyDeviceToHost, c->stream));
  if (rcr) IA_HIP(c, hipMemcpyAsync(rcr, c->rec[2], clen, hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_analyse_picture(ia_ctx *c, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp, int kind,
                       const int *ref_slots, int nref, ia_mbcode *codes, ia_mbstats *stats, uint8_t *ry, uint8_t *rcb,
                       uint8_t *rcr, ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  int rc = ia_upload_source(c, y, cb, cr);
  if (rc) return rc;
  if (kind == IA_PIC_P && c->cfg.search_mode != IA_SEARCH_DIAMOND) {
    rc = ia_motion_search(c, ref_slots, nref);
    if (rc) return rc;
  }
  rc = ia_analyse_uploaded(c, qp, kind, ref_slots, nref);
  if (rc) return rc;
  return ia_read_analysis(c, codes, stats, ry, rcb, rcr, mbs);
}

int ia_sync(ia_ctx *c) {
  if (!c) return IA_EINVAL;
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

static uint16_t f2bf_host(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

int ia_load_proposal_weights(ia_ctx *c, const float *params, int n) {
  if (!c) return IA_EINVAL;
  if (!params || n != IA_CNN_PARAMS) return fail(c, IA_EINVAL, "expected IA_CNN_PARAMS parameters");
  static const int CIN[4] = {8, 32, 32, 32}, COUT[4] = {32, 32, 32, 4};
  ia::CnnWeights *w = new ia::CnnWeights();
  memset(w, 0, sizeof *w);
  const float *q = params;
  for (int l = 0; l < 4; l++) {
    for (int co = 0; co < COUT[l]; co++)
      for (int ci = 0; ci < CIN[l]; ci++)
        for (int t = 0; t < 9; t++) w->w[l][t][co][ci] = f2bf_host(q[(co * CIN[l] + ci) * 9 + t]);
    q += COUT[l] * CIN[l] * 9;
    for (int co = 0; co < COUT[l]; co++) w->b[l][co] = q[co];
    q += COUT[l];
  }
  // FNV-1a of the converted weights: a batch's learned streams must share theirs
  uint64_t h = 1469598103934665603ull;
  const uint8_t *bytes = reinterpret_cast<const uint8_t *>(w);
  for (size_t i = 0; i < sizeof *w; i++) h = (h ^ bytes[i]) * 1099511628211ull;
  hipError_t e = hipMemcpy(c->wts, w, sizeof *w, hipMemcpyHostToDevice);
  delete w;
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(weights)");
  c->wts_loaded = true;
  c->wts_hash = h;
  return IA_OK;
}

static int ensure_tool_buffers(ia_ctx *c) {
  const size_t nmb = (size_t)c->g.mb_w * c->g.mb_h;
  if (!c->feat) IA_HIP(c, hipMalloc((void **)&c->feat, nmb * IA_CNN_CELLS * IA_CNN_FEATURES * sizeof(float)));
  if (!c->cnn_out) IA_HIP(c, hipMalloc((void **)&c->cnn_out, nmb * IA_CNN_CELLS * IA_CNN_OUTPUTS * sizeof(float)));
  if (!c->centres) IA_HIP(c, hipMalloc((void **)&c->centres, nmb * 6 * sizeof(int16_t)));
  return IA_OK;
}

int ia_proposal_features(ia_ctx *c, const int *ref_slots, int nref, float *features, int16_t *centres) {
  if (!c) return IA_EINVAL;
  int rc = iah::check_refs(c, ref_slots, nref);
  if (rc) return rc;
  if (nref > 3) return fail(c, IA_EUNSUPPORTED, "the learned proposal takes at most 3 references");
  rc = ensure_tool_buffers(c);
  if (rc) return rc;
  ia::PropArgs p;
  memset(&p, 0, sizeof p);
  p.src_y = c->src[0];
  for (int r = 0; r < nref; r++) p.ref_g[r] = c->slots[ref_slots[r]].hp;
  p.rmbs0 = c->slots[ref_slots[0]].mbs;
  p.wts = c->wts;
  p.feat = c->feat;
  p.centre = c->centres;
  p.mb_w = c->g.mb_w;
  p.mb_h = c->g.mb_h;
  p.nref = nref;
  p.features_only = 1;
  IA_HIP(c, ia::launch_proposal(p, c->stream));
  const size_t nmb = (size_t)c->g.mb_w * c->g.mb_h;
  if (features)
    IA_HIP(c, hipMemcpyAsync(features, c->feat, nmb * IA_CNN_CELLS * IA_CNN_FEATURES * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (centres) IA_HIP(c, hipMemcpyAsync(centres, c->centres, nmb * 6 * sizeof(int16_t), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_proposal_forward(ia_ctx *c, const float *features, int n, float *outputs) {
  if (!c) return IA_EINVAL;
  const int nmb = c->g.mb_w * c->g.mb_h;
  if (!features || !outputs || n < 1 || n > nmb) return fail(c, IA_EINVAL, "bad features or count");
  if (!c->wts_loaded) return fail(c, IA_ESTATE, "the proposal's weights are not loaded");
  int rc = ensure_tool_buffers(c);
  if (rc) return rc;
  IA_HIP(c, hipMemcpyAsync(c->feat, features, (size_t)n * IA_CNN_CELLS * IA_CNN_FEATURES * sizeof(float), hipMemcpyHostToDevice, c->stream));
  ia::PropArgs p;
  memset(&p, 0, sizeof p);
  p.src_y = c->src[0];
  p.wts = c->wts;
  p.feat_in = c->feat;
  p.cnn_out = c->cnn_out;
  p.mb_w = n;  // one row of n macroblocks: only the count matters with given features
  p.mb_h = 1;
  p.nref = 1;
  IA_HIP(c, ia::launch_proposal(p, c->stream));
  IA_HIP(c, hipMemcpyAsync(outputs, c->cnn_out, (size_t)n * IA_CNN_CELLS * IA_CNN_OUTPUTS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_read_fallback_count(ia_ctx *c, int *count) {
  if (!c || !count) return IA_EINVAL;
  // the count of the last learned search (the other word once a search has flipped them)
  IA_HIP(c, hipMemcpyAsync(count, c->fb + (c->fb_gen ^ 1), sizeof(int), hipMemcpyDeviceToHost, c->stream));
  IA_HIP(c, hipStreamSynchronize(c->stream));
  return IA_OK;
}

int ia_fullsearch_variants(void) { return ia::fullsearch_variants(); }

const char *ia_fullsearch_variant_name(int v) { return ia::fullsearch_variant_name(v); }

int ia_set_fullsearch_variant(ia_ctx *c, int v) {
  if (!c) return IA_EINVAL;
  if (v < -1 || v >= ia::fullsearch_variants()) return fail(c, IA_EINVAL, "unknown variant");
  c->fs_variant = v;
  return IA_OK;
}

}  // extern "C"
