// api.hip — the C ABI of include/ironav_h264a.h: context, device buffers and launches.
#include <stdio.h>
#include <string.h>

#include <string>

#include "common.h"

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
  std::string err;
};

namespace {

thread_local std::string g_create_error;

int fail(ia_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

int hip_fail(ia_ctx *c, hipError_t e, const char *what) {
  return fail(c, IA_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define IA_HIP(c, call)                                    \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
  } while (0)

int check_config(const ia_config *cfg, std::string &why) {
  if (cfg->width < 16 || cfg->height < 16 || cfg->width > 8192 || cfg->height > 8192 || (cfg->width & 1) || (cfg->height & 1)) {
    why = "width and height must be even and within 16..8192";
    return IA_EINVAL;
  }
  if (cfg->bit_depth != 8) {
    why = "only 8-bit pictures are supported on the GPU";
    return IA_EUNSUPPORTED;
  }
  if (cfg->fast) {
    why = "the fast (realtime) decisions are not implemented";
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

const char *ia_last_error(const ia_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

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

static int check_refs(ia_ctx *c, const int *ref_slots, int nref) {
  if (nref < 1 || nref > c->cfg.nrefs || !ref_slots) return fail(c, IA_EINVAL, "nref must be within 1..nrefs");
  for (int r = 0; r < nref; r++) {
    if (ref_slots[r] < 0 || ref_slots[r] >= c->cfg.nrefs) return fail(c, IA_EINVAL, "reference slot out of range");
    if (!c->slots[ref_slots[r]].loaded) return fail(c, IA_ESTATE, "reference slot is empty");
  }
  return IA_OK;
}

int ia_motion_search(ia_ctx *c, const int *ref_slots, int nref) {
  if (!c) return IA_EINVAL;
  int rc = check_refs(c, ref_slots, nref);
  if (rc) return rc;
  ia::FsArgs a;
  memset(&a, 0, sizeof a);
  a.src_y = c->src[0];
  for (int r = 0; r < nref; r++) a.ref_g[r] = c->slots[ref_slots[r]].hp;
  a.out = c->fs;
  a.nref = nref;
  a.mb_w = c->g.mb_w;
  a.mb_h = c->g.mb_h;
  a.range = c->cfg.search_range;
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
  return fail(c, IA_EUNSUPPORTED, "the analysis kernel is not built yet");
}

int ia_read_analysis(ia_ctx *c, ia_mbcode *codes, ia_mbstats *stats, uint8_t *ry, uint8_t *rcb, uint8_t *rcr,
                     ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  return fail(c, IA_EUNSUPPORTED, "the analysis kernel is not built yet");
}

int ia_analyse_picture(ia_ctx *c, const uint8_t *y, const uint8_t *cb, const uint8_t *cr, int qp, int kind,
                       const int *ref_slots, int nref, ia_mbcode *codes, ia_mbstats *stats, uint8_t *ry, uint8_t *rcb,
                       uint8_t *rcr, ia_mbinfo *mbs) {
  if (!c) return IA_EINVAL;
  int rc = ia_upload_source(c, y, cb, cr);
  if (rc) return rc;
  if (kind == IA_PIC_P && c->cfg.search_mode == IA_SEARCH_FULL) {
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

int ia_fullsearch_variants(void) { return ia::fullsearch_variants(); }

const char *ia_fullsearch_variant_name(int v) { return ia::fullsearch_variant_name(v); }

int ia_set_fullsearch_variant(ia_ctx *c, int v) {
  if (!c) return IA_EINVAL;
  if (v < -1 || v >= ia::fullsearch_variants()) return fail(c, IA_EINVAL, "unknown variant");
  c->fs_variant = v;
  return IA_OK;
}

}  // extern "C"
