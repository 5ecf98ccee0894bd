// batch.hip — many analysis contexts (streams of any sizes) through one set of launches per frame
// step, optionally captured once as a hipGraph and replayed.
//
// A step enqueues, on the batch's stream:
//   1. the copy of the per-frame table (AnParams, FsArgs, half-sample and copy jobs per stream) from
//      pinned host memory to the device;
//   2. the reset of the analysis queue (a one-word kernel: no memset nodes in the graph);
//   3. the exhaustive search of every (stream, macroblock, reference), one launch per tile shape
//      (streams of IA_SEARCH_FULL; the autotuner sets each context's shape); for the streams of
//      IA_SEARCH_LEARNED, the proposal over all their macroblocks (one persistent launch; the CNN in
//      the same graph) and the exhaustive search of the batch's one fallback list;
//   4. the analysis: a pool of workgroups taking (stream, slice) tasks from a queue, longest first;
//   5. the promotion of each stream's reconstruction to a reference slot: its planes and macroblock
//      records copied (one batched copy; the luma to the context's scratch plane), the loop filter
//      of the copies (one launch over every promoted stream's macroblock rows, when the frame asks
//      for it, as the encoder filters its references), then the half-sample planes (one launch).
// Every kernel reads its per-stream arguments from the device table, so one captured graph serves
// every frame: the grids cover the largest case (all references of every stream) and the blocks of
// a stream's unused references, or of a stream that promotes nothing, return at once.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.h"

namespace {

struct CopyJob {
  uint8_t *dst;                // null: nothing to copy this step
  const uint8_t *src;
  long long bytes;
  int first, pad;              // first block of the job in the launch
};

constexpr int CP_THREADS = 256, CP_UNITS = 4;                            // 16-byte units per thread
constexpr long long CP_CHUNK = (long long)CP_THREADS * CP_UNITS * 16;  // bytes per block

// Block b copies chunk b - first of the job whose chunks hold it (jobs by `first`).
__global__ __launch_bounds__(CP_THREADS) void copy_batch_kernel(const CopyJob *jobs, int njobs) {
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= b) lo = mid;
    else hi = mid - 1;
  }
  const CopyJob &j = jobs[lo];
  if (!j.dst) return;
  const long long base = (long long)(b - j.first) * CP_CHUNK;
#pragma unroll
  for (int k = 0; k < CP_UNITS; k++) {
    const long long o = base + ((long long)k * CP_THREADS + threadIdx.x) * 16;
    if (o + 16 <= j.bytes) {
      *reinterpret_cast<uint4 *>(j.dst + o) = *reinterpret_cast<const uint4 *>(j.src + o);
    } else {
      for (long long q = o; q < j.bytes && q < o + 16; q++) j.dst[q] = j.src[q];
    }
  }
}

__global__ void reset_words(int *p, int *q) {
  if (threadIdx.x == 0) *p = 0;
  if (threadIdx.x == 1 && q) *q = 0;
}

size_t align16(size_t n) { return (n + 15) & ~(size_t)15; }

}  // namespace

struct ia_batch {
  std::vector<ia_ctx *> ctx;
  int n = 0;
  hipStream_t stream = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the per-frame table, pinned on the host and mirrored on the device
  size_t tab_bytes = 0, off_fs = 0, off_hp = 0, off_cp = 0, off_db = 0, off_pp = 0;
  uint8_t *h_tab = nullptr, *d_tab = nullptr;
  // fixed tables: analysis tasks, queue head, search launch members
  int *d_fixed = nullptr;
  int ntasks = 0, workers = 0, hp_blocks = 0, cp_blocks = 0, db_blocks = 0;
  struct Group {
    int variant, max_range, nmem, blocks;
    size_t offs, sidx;  // word offsets into d_fixed
  };
  std::vector<Group> groups;
  size_t head = 0, tasks = 0;  // word offsets into d_fixed
  // learned streams: proposal launch members (word offsets into d_fixed), the batch's fallback list
  size_t prop_offs = 0, prop_sidx = 0;
  int prop_nmem = 0, prop_total = 0, fb_max_range = 0, fb_max_entries = 0;
  int *d_fb = nullptr;  // [0] count, [1..] entries picture << 20 | mb << 2 | ref
  int *d_dberr = nullptr;  // set when a loop filter row stalled
  uint64_t wts_hash = 0;
  std::string err;

  ia::AnParams *an_h() { return reinterpret_cast<ia::AnParams *>(h_tab); }
  ia::FsArgs *fs_h() { return reinterpret_cast<ia::FsArgs *>(h_tab + off_fs); }
  ia::HpJob *hp_h() { return reinterpret_cast<ia::HpJob *>(h_tab + off_hp); }
  CopyJob *cp_h() { return reinterpret_cast<CopyJob *>(h_tab + off_cp); }
  ia::DbJob *db_h() { return reinterpret_cast<ia::DbJob *>(h_tab + off_db); }
  ia::PropArgs *pp_h() { return reinterpret_cast<ia::PropArgs *>(h_tab + off_pp); }
};

namespace {

int bfail(ia_batch *b, int code, const std::string &msg) {
  b->err = msg;
  return code;
}

#define IA_BHIP(b, call)                                                                       \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return bfail((b), IA_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int enqueue(ia_batch *b) {
  hipStream_t st = b->stream;
  IA_BHIP(b, hipMemcpyAsync(b->d_tab, b->h_tab, b->tab_bytes, hipMemcpyHostToDevice, st));
  int *fixed = b->d_fixed;
  hipLaunchKernelGGL(reset_words, dim3(1), dim3(64), 0, st, fixed + b->head, b->d_fb);
  IA_BHIP(b, hipGetLastError());
  const ia::FsArgs *fs_d = reinterpret_cast<const ia::FsArgs *>(b->d_tab + b->off_fs);
  for (const auto &g : b->groups)
    IA_BHIP(b, ia::launch_fullsearch_batch(fs_d, fixed + g.offs, fixed + g.sidx, g.nmem, g.blocks, g.variant,
                                           g.max_range, st));
  if (b->prop_nmem > 0) {
    IA_BHIP(b, ia::launch_proposal_batch(reinterpret_cast<const ia::PropArgs *>(b->d_tab + b->off_pp),
                                         fixed + b->prop_offs, fixed + b->prop_sidx, b->prop_nmem, b->prop_total, st));
    if (b->fb_max_entries > 0)
      IA_BHIP(b, ia::launch_fullsearch_list_batch(fs_d, b->d_fb + 1, b->d_fb, b->fb_max_range, b->fb_max_entries, st));
  }
  IA_BHIP(b, ia::launch_analyse_batch(reinterpret_cast<const ia::AnParams *>(b->d_tab), fixed + b->tasks, b->ntasks,
                                      fixed + b->head, b->workers, st));
  hipLaunchKernelGGL(copy_batch_kernel, dim3(b->cp_blocks), dim3(CP_THREADS), 0, st,
                     reinterpret_cast<const CopyJob *>(b->d_tab + b->off_cp), 4 * b->n);
  IA_BHIP(b, hipGetLastError());
  IA_BHIP(b, ia::launch_deblock_batch(reinterpret_cast<const ia::DbJob *>(b->d_tab + b->off_db), b->n, b->db_blocks,
                                      b->d_dberr, st));
  IA_BHIP(b, ia::launch_halfpel_batch(reinterpret_cast<const ia::HpJob *>(b->d_tab + b->off_hp), b->n, b->hp_blocks,
                                      st));
  return IA_OK;
}

}  // namespace

extern "C" {

int ia_batch_create(ia_batch **out, ia_ctx *const *ctxs, int n, int use_graph) {
  if (!out || !ctxs || n < 1 || n > 65535) return IA_EINVAL;
  *out = nullptr;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) return IA_EINVAL;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return IA_EINVAL;
  }
  if (n > 2047) {
    for (int i = 0; i < n; i++)
      if (ctxs[i]->cfg.search_mode == IA_SEARCH_LEARNED) return IA_EINVAL;  // fallback entries tag 11 bits
  }
  ia_batch *b = new ia_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->n = n;
  b->use_graph = use_graph != 0;
  // the per-frame table
  b->off_fs = align16((size_t)n * sizeof(ia::AnParams));
  b->off_hp = b->off_fs + align16((size_t)n * sizeof(ia::FsArgs));
  b->off_cp = b->off_hp + align16((size_t)n * sizeof(ia::HpJob));
  b->off_db = b->off_cp + align16((size_t)4 * n * sizeof(CopyJob));
  b->off_pp = b->off_db + align16((size_t)n * sizeof(ia::DbJob));
  b->tab_bytes = b->off_pp + align16((size_t)n * sizeof(ia::PropArgs));
  // analysis tasks, longest slice first (ties by stream, then slice)
  struct Task {
    int mbs, s, k;
  };
  std::vector<Task> tv;
  for (int s = 0; s < n; s++) {
    const ia_ctx *c = ctxs[s];
    const int ns = std::min(c->cfg.slices, c->g.mb_h);
    for (int k = 0; k < ns; k++)
      tv.push_back({((k + 1) * c->g.mb_h / ns - k * c->g.mb_h / ns) * c->g.mb_w, s, k});
  }
  std::stable_sort(tv.begin(), tv.end(), [](const Task &x, const Task &y) { return x.mbs > y.mbs; });
  std::vector<int> fixed;
  b->head = fixed.size();
  fixed.push_back(0);
  b->tasks = fixed.size();
  for (const Task &t : tv) fixed.push_back(t.s << 16 | t.k);
  b->ntasks = (int)tv.size();
  // the exhaustive search: one launch per (tile shape, window class) of the full-search streams
  for (int s = 0; s < n; s++) {
    const ia_ctx *c = ctxs[s];
    if (c->cfg.search_mode != IA_SEARCH_FULL) continue;
    const int v = c->fs_variant >= 0 && c->fs_variant < ia::fullsearch_variants() ? c->fs_variant
                                                                                   : ia::fullsearch_default_variant();
    const int wide = c->cfg.search_range > 32;
    ia_batch::Group *g = nullptr;
    for (auto &x : b->groups)
      if (x.variant == v && (x.max_range > 32) == (wide != 0)) g = &x;
    if (!g) {
      b->groups.push_back({v, 0, 0, 0, 0, 0});
      g = &b->groups.back();
    }
    g->max_range = std::max(g->max_range, c->cfg.search_range);
  }
  for (auto &g : b->groups) {
    std::vector<int> offs, sidx;
    int blocks = 0;
    for (int s = 0; s < n; s++) {
      const ia_ctx *c = ctxs[s];
      if (c->cfg.search_mode != IA_SEARCH_FULL) continue;
      const int v = c->fs_variant >= 0 && c->fs_variant < ia::fullsearch_variants() ? c->fs_variant
                                                                                     : ia::fullsearch_default_variant();
      if (v != g.variant || (c->cfg.search_range > 32) != (g.max_range > 32)) continue;
      const int nmb = c->g.mb_w * c->g.mb_h, nmbp = (nmb + 7) / 8 * 8;
      offs.push_back(blocks);
      sidx.push_back(s);
      blocks += nmbp * c->cfg.nrefs;
    }
    g.nmem = (int)offs.size();
    g.blocks = blocks;
    g.offs = fixed.size();
    fixed.insert(fixed.end(), offs.begin(), offs.end());
    g.sidx = fixed.size();
    fixed.insert(fixed.end(), sidx.begin(), sidx.end());
  }
  // the learned streams: one proposal launch over all their macroblocks; the fallback list holds at
  // most every (macroblock, reference) of them
  {
    std::vector<int> offs, sidx;
    int total = 0;
    for (int s = 0; s < n; s++) {
      const ia_ctx *c = ctxs[s];
      if (c->cfg.search_mode != IA_SEARCH_LEARNED) continue;
      offs.push_back(total);
      sidx.push_back(s);
      total += c->g.mb_w * c->g.mb_h;
      b->fb_max_range = std::max(b->fb_max_range, c->cfg.search_range);
      if (std::isfinite(c->cfg.fallback_ratio)) b->fb_max_entries += c->g.mb_w * c->g.mb_h * std::min(c->cfg.nrefs, 3);
    }
    b->prop_nmem = (int)offs.size();
    b->prop_total = total;
    b->prop_offs = fixed.size();
    fixed.insert(fixed.end(), offs.begin(), offs.end());
    b->prop_sidx = fixed.size();
    fixed.insert(fixed.end(), sidx.begin(), sidx.end());
  }
  // promotion launches: strips of each stream's half-sample planes, chunks of its copies, rows of
  // its loop filter
  int hp_blocks = 0, cp_blocks = 0, db_blocks = 0;
  for (int s = 0; s < n; s++) {
    const Geom &g = ctxs[s]->g;
    hp_blocks += ((g.sw + ia::HP_TW - 1) / ia::HP_TW) * ((g.sh + ia::HP_TH - 1) / ia::HP_TH);
    const long long ylen = (long long)g.pw * g.ph, clen = (long long)g.cw * g.ch,
                    mlen = (long long)g.mb_w * g.mb_h * sizeof(ia_mbinfo);
    cp_blocks += (int)((ylen + CP_CHUNK - 1) / CP_CHUNK) + 2 * (int)((clen + CP_CHUNK - 1) / CP_CHUNK) +
                 (int)((mlen + CP_CHUNK - 1) / CP_CHUNK);
    db_blocks += g.mb_h;
  }
  b->hp_blocks = hp_blocks;
  b->cp_blocks = cp_blocks;
  b->db_blocks = db_blocks;
  const int workers = ia::analyse_batch_workers();
  b->workers = std::max(1, std::min(workers > 0 ? workers : 1024, b->ntasks));
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipHostMalloc((void **)&b->h_tab, b->tab_bytes, hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void **)&b->d_tab, b->tab_bytes) == hipSuccess &&
            hipMalloc((void **)&b->d_fixed, fixed.size() * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&b->d_fb, (1 + (size_t)b->fb_max_entries) * sizeof(int)) == hipSuccess &&
            hipMalloc((void **)&b->d_dberr, sizeof(int)) == hipSuccess &&
            hipMemset(b->d_dberr, 0, sizeof(int)) == hipSuccess &&
            hipMemcpy(b->d_fixed, fixed.data(), fixed.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    ia_batch_destroy(b);
    return IA_ENOMEM;
  }
  memset(b->h_tab, 0, b->tab_bytes);
  *out = b;
  return IA_OK;
}

void ia_batch_destroy(ia_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  (void)hipFree(b->d_tab);
  (void)hipFree(b->d_fixed);
  (void)hipFree(b->d_fb);
  (void)hipFree(b->d_dberr);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *ia_batch_last_error(const ia_batch *b) { return b ? b->err.c_str() : "null batch"; }

int ia_batch_step(ia_batch *b, const ia_frame *frames) {
  if (!b || !frames) return IA_EINVAL;
  // the previous step has read the table; the contexts' own streams (uploads) are done
  IA_BHIP(b, hipStreamSynchronize(b->stream));
  for (ia_ctx *c : b->ctx) IA_BHIP(b, hipStreamSynchronize(c->stream));
  for (int s = 0; s < b->n; s++) {
    ia_ctx *c = b->ctx[s];
    const ia_frame &f = frames[s];
    if (f.promote_slot < -1 || f.promote_slot >= c->cfg.nrefs)
      return bfail(b, IA_EINVAL, "stream " + std::to_string(s) + ": promote_slot out of range");
    const uint8_t *src[3] = {f.y, f.cb, f.cr};
    // the searches run for P pictures only (a B picture searches inside the analysis)
    const int nref = f.kind == IA_PIC_P ? f.nref : 0;
    int rc;
    if (f.kind == IA_PIC_B) {
      if (f.nref != 2) return bfail(b, IA_EINVAL, "stream " + std::to_string(s) + ": a B picture takes one picture per list (nref 2)");
      if (f.promote_slot >= 0) return bfail(b, IA_EINVAL, "stream " + std::to_string(s) + ": a B picture is not a reference");
      rc = iah::b_params(c, f.qp, f.ref_slots[0], f.ref_slots[1], f.dist_l0, f.dist_l1, src, &b->an_h()[s]);
    } else {
      rc = iah::analysis_params(c, f.qp, f.kind, f.ref_slots, nref, src, false, &b->an_h()[s]);
    }
    if (rc) return bfail(b, rc, "stream " + std::to_string(s) + ": " + c->err);
    ia::FsArgs &fa = b->fs_h()[s];
    ia::PropArgs &pa = b->pp_h()[s];
    memset(&pa, 0, sizeof pa);
    if (c->cfg.search_mode == IA_SEARCH_LEARNED) {
      if (!c->wts_loaded) return bfail(b, IA_ESTATE, "stream " + std::to_string(s) + ": the proposal's weights are not loaded");
      if (nref > 3) return bfail(b, IA_EUNSUPPORTED, "stream " + std::to_string(s) + ": the learned proposal takes at most 3 references");
      if (b->wts_hash == 0) b->wts_hash = c->wts_hash;
      if (c->wts_hash != b->wts_hash) return bfail(b, IA_ESTATE, "the learned streams of a batch must share their CNN weights");
      pa.src_y = f.y ? f.y : c->src[0];
      for (int r = 0; r < nref; r++) pa.ref_g[r] = c->slots[f.ref_slots[r]].hp;
      pa.rmbs0 = nref > 0 ? c->slots[f.ref_slots[0]].mbs : nullptr;
      pa.wts = c->wts;
      pa.out = c->fs;
      pa.fb_count = b->d_fb;
      pa.fb_list = std::isfinite(c->cfg.fallback_ratio) ? b->d_fb + 1 : nullptr;
      pa.fb_tag = s << 20;
      pa.mb_w = c->g.mb_w;
      pa.mb_h = c->g.mb_h;
      pa.nref = nref;
      pa.fallback_ratio = c->cfg.fallback_ratio;
    }
    if (c->cfg.search_mode != IA_SEARCH_DIAMOND) {
      iah::search_args(c, f.ref_slots, nref, f.y, &fa);
    } else {
      memset(&fa, 0, sizeof fa);
      fa.mb_w = c->g.mb_w;
      fa.mb_h = c->g.mb_h;
    }
    const Geom &g = c->g;
    IaSlot *ps = f.promote_slot >= 0 ? &c->slots[f.promote_slot] : nullptr;
    // the luma goes through the context's scratch plane (filtered there), the rest to the slot
    ia::HpJob &hj = b->hp_h()[s];
    hj.luma = c->tmp_luma;
    hj.planes = ps ? ps->hp : nullptr;
    hj.pw = g.pw;
    hj.ph = g.ph;
    hj.tiles_x = (g.sw + ia::HP_TW - 1) / ia::HP_TW;
    CopyJob *cj = b->cp_h() + 4 * s;
    const long long ylen = (long long)g.pw * g.ph, clen = (long long)g.cw * g.ch,
                    mlen = (long long)g.mb_w * g.mb_h * sizeof(ia_mbinfo);
    cj[0] = {ps ? c->tmp_luma : nullptr, c->rec[0], ylen, 0, 0};
    cj[1] = {ps ? ps->cb : nullptr, c->rec[1], clen, 0, 0};
    cj[2] = {ps ? ps->cr : nullptr, c->rec[2], clen, 0, 0};
    cj[3] = {ps ? reinterpret_cast<uint8_t *>(ps->mbs) : nullptr, reinterpret_cast<const uint8_t *>(c->mbs), mlen, 0, 0};
    ia::DbJob &dj = b->db_h()[s];
    const bool filt = ps && f.promote_deblock;
    dj.y = filt ? c->tmp_luma : nullptr;
    dj.cb = filt ? ps->cb : nullptr;
    dj.cr = filt ? ps->cr : nullptr;
    dj.mbs = c->mbs;
    dj.prog = c->db + 1;
    dj.mb_w = g.mb_w;
    dj.mb_h = g.mb_h;
  }
  // first blocks of the promotion jobs (the same every step)
  int hfirst = 0, cfirst = 0, dfirst = 0;
  for (int s = 0; s < b->n; s++) {
    const Geom &g = b->ctx[s]->g;
    b->hp_h()[s].first = hfirst;
    hfirst += ((g.sw + ia::HP_TW - 1) / ia::HP_TW) * ((g.sh + ia::HP_TH - 1) / ia::HP_TH);
    b->db_h()[s].first = dfirst;
    dfirst += g.mb_h;
    for (int k = 0; k < 4; k++) {
      CopyJob &j = b->cp_h()[4 * s + k];
      j.first = cfirst;
      cfirst += (int)((j.bytes + CP_CHUNK - 1) / CP_CHUNK);
    }
  }
  if (b->use_graph) {
    if (!b->exec) {
      IA_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue(b);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return bfail(b, IA_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      b->graph = graph;
      IA_BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    IA_BHIP(b, hipGraphLaunch(b->exec, b->stream));
  } else {
    int rc = enqueue(b);
    if (rc) return rc;
  }
  // the contexts' state after the step
  for (int s = 0; s < b->n; s++) {
    ia_ctx *c = b->ctx[s];
    const ia_frame &f = frames[s];
    const int nref = f.kind == IA_PIC_P ? f.nref : 0;
    if (c->cfg.search_mode != IA_SEARCH_DIAMOND && nref > 0) {
      c->fs_nref = nref;
      for (int r = 0; r < nref; r++) c->fs_slots[r] = f.ref_slots[r];
      c->fs_valid = true;
    } else {
      c->fs_valid = false;
    }
    c->an_valid = true;
    c->an_b = f.kind == IA_PIC_B;
    if (f.promote_slot >= 0) c->slots[f.promote_slot].loaded = true;
  }
  return IA_OK;
}

int ia_batch_fallback_count(ia_batch *b, int *count) {
  if (!b || !count) return IA_EINVAL;
  IA_BHIP(b, hipStreamSynchronize(b->stream));
  *count = 0;
  if (b->d_fb) IA_BHIP(b, hipMemcpy(count, b->d_fb, sizeof(int), hipMemcpyDeviceToHost));
  return IA_OK;
}

int ia_batch_workers(void) { return ia::analyse_batch_workers(); }

int ia_batch_sync(ia_batch *b) {
  if (!b) return IA_EINVAL;
  IA_BHIP(b, hipStreamSynchronize(b->stream));
  int stalled = 0;
  IA_BHIP(b, hipMemcpy(&stalled, b->d_dberr, sizeof(int), hipMemcpyDeviceToHost));
  if (stalled) {
    // A row that gave up waiting cleared the counter of the row above it, which may have published
    // later: reset every stream's row counters (outside the graph, the step is over) so that the next
    // step's rows wait as they should, as the single-context promotion does after a stall.
    IA_BHIP(b, hipMemset(b->d_dberr, 0, sizeof(int)));
    for (ia_ctx *c : b->ctx) IA_BHIP(b, hipMemset(c->db, 0, (1 + (size_t)c->g.mb_h) * sizeof(int)));
    return bfail(b, IA_EDEVICE, "the loop filter's rows stalled");
  }
  return IA_OK;
}

}  // extern "C"
