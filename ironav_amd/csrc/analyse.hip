// analyse.hip — the per-macroblock analysis of H.264 P and I slices on gfx950 (code_mb and what it calls).
//
// Reproduces nervosys/IronAV's slice coder decisions (crates/ironav-codec/src/h264/enc/frame.rs):
// P_Skip test, Intra16x16 mode choice, motion search and partition decision over every list-0
// reference, the rate-distortion choice among P_Skip, inter (4x4 and 8x8 transform), Intra16x16,
// Intra4x4 and Intra8x8 with the CABAC bit model, the transforms, quantisation and reconstruction,
// and the neighbour state and CABAC context evolution from one macroblock to the next.
//
// A slice's macroblocks depend on their left, top and top-right neighbours' final decisions and
// reconstruction, and on the CABAC contexts the previous macroblock left, so each slice is one
// workgroup walking its macroblocks in raster order; the parallelism inside a macroblock is across
// the lanes of a wave (pixels, 4x4 blocks, prediction modes). All lanes run the decision logic in
// lock step on the same values; loops over samples and blocks are split across lanes and summed
// with DPP reductions.
//
// The macroblock being coded lives in an LDS canvas together with the samples its intra
// predictions read (the row above including the above-right macroblock, the column to the left):
// candidates are coded into the canvas as the reference codes them into its picture, and the kept
// one is written to the picture at the end. Decision arithmetic in f64 (the skip threshold, lam2,
// J = SSD + lam2 * bits / 256) is computed without contraction (the file is built with
// -ffp-contract=off), with the powers of two computed on the host by the C library as the
// reference computes them.
#include <cstddef>

#include "analyse.h"

// Pointers qualified with it are global memory: their loads are global_load instructions, which wait
// on the vector memory counter alone (generic pointers' FLAT loads wait on the LDS counter too).
#ifndef IA_GLOBAL_AS
#define IA_GLOBAL_AS __attribute__((address_space(1)))
#endif
// a generic pointer to global memory as such (a pointer argument, so that no integer converts)
template <class T>
__device__ __forceinline__ IA_GLOBAL_AS const T *as_global(const T *p) {
  return (IA_GLOBAL_AS const T *)p;
}

namespace ia {
namespace {

enum { K_NONE = 0, K_I4 = 1, K_I8 = 2, K_I16 = 3, K_INTER = 5 };

// ---- tables --------------------------------------------------------------------------------------

// quarter-sample fraction -> (plane a, dx, dy, plane b, dx, dy), indexed [fy][fx]; planes 0 G, 1 B,
// 2 H, 3 J: the two integer/half samples H.264 8.4.2.2.1 averages (enc/subpel.rs:26-38)
constexpr uint8_t QMIX[4][4][6] = {
    {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 0}, {1, 0, 0, 1, 0, 0}, {0, 1, 0, 1, 0, 0}},
    {{0, 0, 0, 2, 0, 0}, {1, 0, 0, 2, 0, 0}, {1, 0, 0, 3, 0, 0}, {1, 0, 0, 2, 1, 0}},
    {{2, 0, 0, 2, 0, 0}, {2, 0, 0, 3, 0, 0}, {3, 0, 0, 3, 0, 0}, {3, 0, 0, 2, 1, 0}},
    {{0, 0, 1, 2, 0, 0}, {2, 0, 0, 1, 0, 1}, {3, 0, 0, 1, 0, 1}, {2, 1, 0, 1, 0, 1}},
};
// QMIX in registers: the entries of phase e = 4 fy + fx a byte each (plane a in bits 0-1, its column
// and row offsets in bits 2 and 3, plane b in bits 4-5, its offsets in bits 6 and 7), eight to a word,
// so a lane's phase reads no table from memory
constexpr uint64_t qmix_pack(int half) {
  uint64_t k = 0;
  for (int e = 0; e < 8; e++) {
    const uint8_t *m = QMIX[(e + 8 * half) >> 2][(e + 8 * half) & 3];
    k |= (uint64_t)(m[0] | m[1] << 2 | m[2] << 3 | m[3] << 4 | m[4] << 6 | m[5] << 7) << (8 * e);
  }
  return k;
}
constexpr uint64_t QMIX_LO = qmix_pack(0), QMIX_HI = qmix_pack(1);
__device__ __forceinline__ uint32_t qmix(int fy, int fx) {
  const int e = fy * 4 + fx;
  return (uint32_t)(((e < 8 ? QMIX_LO : QMIX_HI) >> (8 * (e & 7))) & 0xffu);
}

__constant__ uint8_t PARTS[4][4][4] = {
    {{0, 0, 4, 4}}, {{0, 0, 4, 2}, {0, 2, 4, 2}}, {{0, 0, 2, 4}, {2, 0, 2, 4}},
    {{0, 0, 2, 2}, {2, 0, 2, 2}, {0, 2, 2, 2}, {2, 2, 2, 2}}};
__constant__ int NPARTS[4] = {1, 2, 2, 4};
__constant__ uint8_t SUBS[4][4][4] = {
    {{0, 0, 2, 2}}, {{0, 0, 2, 1}, {0, 1, 2, 1}}, {{0, 0, 1, 2}, {1, 0, 1, 2}},
    {{0, 0, 1, 1}, {1, 0, 1, 1}, {0, 1, 1, 1}, {1, 1, 1, 1}}};
__constant__ int NSUBS[4] = {1, 2, 2, 4};

#define IA_ZIG4 0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15
__constant__ uint8_t ZIG4[16] = {IA_ZIG4};
// the scan position of each raster position, four bits each
constexpr uint64_t izig4() {
  constexpr uint8_t z[16] = {IA_ZIG4};
  uint64_t k = 0;
  for (int s = 0; s < 16; s++) k |= (uint64_t)s << (4 * z[s]);
  return k;
}
constexpr uint64_t IZIG4 = izig4();
#undef IA_ZIG4
// the frequency each position k of a lane-parallel transform dimension holds (fwd4, inv4): 0, 2, 3, 1
constexpr int perm4(int k) { return (0x1320 >> (4 * k)) & 15; }
__constant__ uint8_t ZIG8[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                                 41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                                 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Rect {
  int x, y, w, h;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }
// v << s on the two's-complement bits (Rust's i32 `<<`, which the dequantisation of negative levels uses)
__device__ __forceinline__ int shl(int v, int s) { return (int)((uint32_t)v << s); }

// ---- wave reductions -------------------------------------------------------------------------------

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ int dpp_add(int v) {
  return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, BANK_MASK, false);
}

// Sum over the 64 lanes, the same in every lane: a Hillis-Steele scan within each row of 16 lanes
// (shifts 1, 2, 4, 8), then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then
// rows 0-1's into rows 2-3 (row_bcast:31); the last lane holds the total. (Each shift adds the
// partial sums once, unlike the 1, 2, 3 shift sequence that suits only idempotent operations.)
__device__ __forceinline__ int wave_sum(int v) {
  v = dpp_add<0x111, 0xf, 0xf>(v);
  v = dpp_add<0x112, 0xf, 0xf>(v);
  v = dpp_add<0x114, 0xf, 0xf>(v);
  v = dpp_add<0x118, 0xf, 0xf>(v);
  v = dpp_add<0x142, 0xa, 0xf>(v);
  v = dpp_add<0x143, 0xc, 0xf>(v);
  return __builtin_amdgcn_readlane(v, 63);
}

// Sum over each group of 4 consecutive lanes, the same in the group's lanes (DPP quad permutations).
__device__ __forceinline__ int quad_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);  // quad_perm [1, 0, 3, 2]
  v += __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);  // quad_perm [2, 3, 0, 1]
  return v;
}

// Sum over each aligned group of n lanes (n = 1, 2, 4, 8 or 16), the same in the group's lanes: the
// quad permutations exchange lanes 1 and 2 apart, the half-row mirror pairs each lane with one of the
// other quad of its 8, the row mirror with one of the other 8 of its 16.
__device__ __forceinline__ int group_sum(int v, int n) {
  if (n >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);   // quad_perm [1, 0, 3, 2]
  if (n >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);   // quad_perm [2, 3, 0, 1]
  if (n >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false);  // row_half_mirror
  if (n >= 16) v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false); // row_mirror
  return v;
}

// lanes whose predicate holds, counted (the same in every lane)
__device__ __forceinline__ int count_lanes(bool pred) { return __popcll(__ballot(pred)); }

__device__ __forceinline__ int lane() { return threadIdx.x; }

__device__ __forceinline__ void sync() { __syncthreads(); }

// ---- exp-Golomb lengths and the MV cost (enc/bitw.rs:81-89, enc/frame.rs:164-171) -----------------

__device__ __forceinline__ int ue_bits(uint32_t v) { return 2 * (32 - __clz((int)(v + 1))) - 1; }
// the signed mapping 2|v| - (v > 0) written without a select (as a select of two expressions the
// compiler branched on it)
__device__ __forceinline__ int se_bits(int v) { return ue_bits(2u * (uint32_t)(v < 0 ? -v : v) - (uint32_t)(v > 0)); }
__device__ __forceinline__ int mvd_bits(int mx, int my, int px, int py) { return se_bits(mx - px) + se_bits(my - py); }

__device__ __forceinline__ int16_t median3(int16_t a, int16_t b, int16_t c) {
  int16_t mx = a > b ? a : b, mn = a < b ? a : b;
  int16_t t = mn > c ? mn : c;
  return mx < t ? mx : t;
}

// ---- per-slice working set in LDS --------------------------------------------------------------

// The result of one reference's search, and the motion search's state per group of 16 lanes: the
// sub-partition search's vectors and references (enc/frame.rs:866) and each group's result.
struct SearchRes {
  int32_t cost, part;
  int16_t mvs[16][2];
  uint8_t refs[4], sub[4];
};
struct SearchLds {
  int16_t st_mvs[4][16][2];
  int8_t st_refs[4][16];
  int16_t tr_mvs[4][16][2];  // the state under trial by the sub-partition search
  int8_t tr_refs[4][16];
  SearchRes res[4];
};
// A B macroblock's working state: the list 1 prediction (averaged with list 0's in L.pred), each
// group's search result, the SATD of each 4x4 block of a prediction.
struct BLds {
  uint8_t p1[256];
  uint8_t c1[2][64];
  int16_t mv[4][2];
  int32_t cost[4];
  int32_t sat[16];
};

// A candidate's chroma coding (mode, nonzero counts, levels, coded pattern), which the rate-distortion
// decision codes once for the intra candidates; the nonzero counts and levels as consecutive words.
struct ChromaSave {
  uint8_t cnnz[2][4];
  int16_t cdc[2][4];
  int16_t cac[2][4][16];
  uint8_t mode, cbp;
};
static_assert(offsetof(ChromaSave, cdc) == 8 && offsetof(ChromaSave, cac) == 24, "the chroma save's words");

struct Lds {
  uint8_t src_y[256];
  uint8_t src_c[2][64];
  // the macroblock and the samples its intra predictions read: luma rows -1..15, columns -1..23
  // (columns 16..23 of row -1 are the above-right macroblock), chroma rows -1..7, columns -1..7
  uint8_t cy[17][28];
  uint8_t cc[2][9][12];
  uint8_t cabac[1024];  // live context states of the slice
  union {
    uint8_t tmp[1024];  // a copy counted on for a rate estimate
    SearchLds ms;       // the motion search's per-group state (not live at the same time)
    BLds bs;            // a B macroblock's predictions and search results (not live at the same time)
    struct {            // the finished macroblock's outputs, formed by one lane and stored by all
      MbState st;
      ia_mbinfo info;
      ia_mbstats stats;
    } fin;
  };
  uint32_t cab[256];    // per (state, LPS?): bit cost in 1/256 bits << 8 | next state (enc/cabac.rs:32-40,162-173)
  MbState nb[4];        // left, top, top-right, top-left neighbours (valid when nb_ok)
  int nb_ok[4];
  ia_mbcode code;       // the candidate being coded
  union {
    ia_mbcode keep;     // the candidate kept by the rate-distortion decision (CABAC)
    ia_mbcode i4;       // Intra4x4 coded ahead of the SATD decision (CAVLC)
  };
  ChromaSave csave;     // the chroma coding the intra candidates share (CABAC)
  uint8_t keep_rec[384];
  uint8_t pred[256];
  uint8_t cpred[2][64];
  int32_t co[16][16];   // dequantised coefficients per 4x4 block (raster)
  int32_t co8[4][64];
  int32_t dcs[16];
  int32_t score[16];
  int32_t ival[64];
  int16_t mv8[16][2];
  uint64_t nzm[2][16];  // masks of the nonzero levels and of the levels above 1 (count_bins)
  uint64_t pos8[24];    // 8x8 scan positions of each significance (0-14) and last (15-23) context
  int etop[16], eleft[8], etl;  // the edges of the Intra4x4 block being coded
  uint32_t kept_bits;           // the kept candidate's rate
  uint8_t kept_st[440];         // the context states (0-439, those the slice codes on) after its bins
  uint8_t chroma_st[64];        // the chroma residual's context states after the last full count
};

// The workgroup's working set. Declared at namespace scope so that every function that touches it
// addresses it as LDS (ds_read/ds_write) rather than through generic pointers.
__shared__ Lds L;

// ---- the slice coder ---------------------------------------------------------------------------

struct Coder {
  const AnParams &P;
  int mbx, mby, row0, row1, slice_id;
  int rd_kept;  // the macroblock was decided by rate-distortion: its rate and states are in L.kept_*

  __device__ Coder(const AnParams &p) : P(p) {}

  __device__ int ls() const { return P.mb_w * 16; }
  __device__ int cs() const { return P.mb_w * 8; }
  __device__ int up() const { return mby > row0; }

  // canvas accessors (x, y relative to the macroblock; -1 is the neighbour row/column)
  __device__ uint8_t &cyv(int x, int y) { return L.cy[y + 1][x + 1]; }
  __device__ uint8_t &ccv(int c, int x, int y) { return L.cc[c][y + 1][x + 1]; }

  // ---- neighbours (enc/frame.rs:509-603) ----

  // neighbour index of macroblock (mbx+dx, mby+dy): 0 left, 1 top, 2 top-right, 3 top-left
  __device__ const MbState *nb(int dx, int dy) const {
    int k = dy == 0 ? 0 : dx == 0 ? 1 : dx > 0 ? 2 : 3;
    return L.nb_ok[k] ? &L.nb[k] : nullptr;
  }

  // list l's reference index and vector of the 4x4 block (x, y) relative to the macroblock
  // (enc/frame.rs:521-539); 0 when it is not available
  __device__ int nb_motion(const int16_t (*cur)[2], const int8_t *cref, uint32_t done, int x, int y, int *r,
                           int16_t *mv, int l = 0) const {
    int dx, dy, bx, by;
    if (x < 0 && y < 0) { dx = -1; dy = -1; bx = 3; by = 3; }
    else if (x < 0) { dx = -1; dy = 0; bx = 3; by = y; }
    else if (x >= 4 && y < 0) { dx = 1; dy = -1; bx = 0; by = 3; }
    else if (x >= 4) return 0;
    else if (y < 0) { dx = 0; dy = -1; bx = x; by = 3; }
    else {
      int b = y * 4 + x;
      if (!((done >> b) & 1)) return 0;
      if ((done >> (16 + b)) & 1) { *r = -1; mv[0] = mv[1] = 0; }
      else { *r = cref[b]; mv[0] = cur[b][0]; mv[1] = cur[b][1]; }
      return 1;
    }
    const MbState *s = nb(dx, dy);
    if (!s) return 0;
    // the list's arrays through pointers (g++ 11 with -fsanitize=undefined, the host build that checks
    // this kernel, subscripts `(l ? a : b)[i]` of two arrays wrongly)
    const int8_t *refs = l ? s->refs1 : s->refs;
    const int16_t(*mvs)[2] = l ? s->mvs1 : s->mvs;
    int rr = refs[(by / 2) * 2 + bx / 2];
    const int16_t *m = mvs[by * 4 + bx];
    if (s->inter && rr >= 0) { *r = rr; mv[0] = m[0]; mv[1] = m[1]; }
    else { *r = -1; mv[0] = mv[1] = 0; }
    return 1;
  }

  __device__ void pred_mv(const int16_t (*cur)[2], const int8_t *cref, uint32_t done, int bx, int by, int w, int dir,
                          int rf, int16_t *out, int l = 0) const {
    int ra = -1, rb = -1, rc = -1;
    int16_t ma[2] = {0, 0}, mb[2] = {0, 0}, mc[2] = {0, 0};
    int ha = nb_motion(cur, cref, done, bx - 1, by, &ra, ma, l);
    int hb = nb_motion(cur, cref, done, bx, by - 1, &rb, mb, l);
    int hc = nb_motion(cur, cref, done, bx + w, by - 1, &rc, mc, l);
    if (!hc) hc = nb_motion(cur, cref, done, bx - 1, by - 1, &rc, mc, l);
    if (!ha) { ra = -1; ma[0] = ma[1] = 0; }
    if (!hb) { rb = -1; mb[0] = mb[1] = 0; }
    if (!hc) { rc = -1; mc[0] = mc[1] = 0; }
    if (dir == 1) {
      if (by == 0 && rb == rf) { out[0] = mb[0]; out[1] = mb[1]; return; }
      if (by != 0 && ra == rf) { out[0] = ma[0]; out[1] = ma[1]; return; }
    } else if (dir == 2) {
      if (bx == 0 && ra == rf) { out[0] = ma[0]; out[1] = ma[1]; return; }
      if (bx != 0 && rc == rf) { out[0] = mc[0]; out[1] = mc[1]; return; }
    }
    if (!hb && !hc && ha) { out[0] = ma[0]; out[1] = ma[1]; return; }
    int ea = ra == rf, eb = rb == rf, ec = rc == rf;
    if (ea + eb + ec == 1) {
      const int16_t *m = ea ? ma : eb ? mb : mc;
      out[0] = m[0];
      out[1] = m[1];
      return;
    }
    out[0] = median3(ma[0], mb[0], mc[0]);
    out[1] = median3(ma[1], mb[1], mc[1]);
  }

  __device__ void mv_pred(int rf, int16_t *out, int l = 0) const {
    const int16_t z[16][2] = {};
    const int8_t nr[16] = {};
    pred_mv(z, nr, 0, 0, 0, 4, 0, rf, out, l);
  }

  __device__ static int rects(int part, const uint8_t *sub, Rect *out) {
    int n = 0;
    if (part != 3) {
      for (int k = 0; k < NPARTS[part]; k++) out[n++] = Rect{PARTS[part][k][0], PARTS[part][k][1], PARTS[part][k][2], PARTS[part][k][3]};
      return n;
    }
    for (int k = 0; k < 4; k++)
      for (int s = 0; s < NSUBS[sub[k]]; s++) {
        const uint8_t *r = SUBS[sub[k]][s];
        out[n++] = Rect{(k & 1) * 2 + r[0], (k >> 1) * 2 + r[1], r[2], r[3]};
      }
    return n;
  }

  __device__ static int ref_of(int part, int n, int x, int y) { return part == 3 ? (y / 2) * 2 + x / 2 : n; }

  __device__ void part_mvps(int part, const uint8_t *sub, const int16_t (*mvs)[2], const uint8_t *refs,
                            int16_t (*out)[2]) const {
    int16_t cur[16][2] = {};
    int8_t cref[16] = {};
    uint32_t done = 0;
    Rect rs[16];
    int n = rects(part, sub, rs);
    const int dir = part == 1 ? 1 : part == 2 ? 2 : 0;
    for (int k = 0; k < n; k++) {
      int rf = refs[ref_of(part, k, rs[k].x, rs[k].y)];
      pred_mv(cur, cref, done, rs[k].x, rs[k].y, rs[k].w, dir, rf, out[k]);
      for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
        for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++) {
          cur[by * 4 + bx][0] = mvs[k][0];
          cur[by * 4 + bx][1] = mvs[k][1];
          cref[by * 4 + bx] = (int8_t)rf;
          done |= 1u << (by * 4 + bx);
        }
    }
  }

  __device__ void skip_mv(int16_t *out) const {
    const int16_t z[16][2] = {};
    const int8_t nr[16] = {};
    int ra, rb;
    int16_t ma[2], mb[2];
    int ha = nb_motion(z, nr, 0, -1, 0, &ra, ma);
    int hb = nb_motion(z, nr, 0, 0, -1, &rb, mb);
    if (!ha || !hb || (ra == 0 && ma[0] == 0 && ma[1] == 0) || (rb == 0 && mb[0] == 0 && mb[1] == 0)) {
      out[0] = out[1] = 0;
      return;
    }
    mv_pred(0, out);
  }

  __device__ int pred_i4_mode(const uint8_t *cur, int bx, int by) const {
    int a = -1, b = -1;
    if (bx > 0) a = cur[by * 4 + bx - 1];
    else if (const MbState *s = nb(-1, 0)) a = s->i4 ? s->modes[by * 4 + 3] : 2;
    if (by > 0) b = cur[(by - 1) * 4 + bx];
    else if (const MbState *s = nb(0, -1)) b = s->i4 ? s->modes[12 + bx] : 2;
    if (a >= 0 && b >= 0) return a < b ? a : b;
    return 2;
  }

  // ---- samples ----

  // ---- motion search: a reference per group of 16 lanes ----
  //
  // best_inter searches its references independently (enc/frame.rs:840-848), so up to four searches
  // run at once, reference c0 + g on lanes 16 g .. 16 g + 15 (the group g). Every value the search
  // code holds is the same within a group; its reductions stay within the group's DPP row and its LDS
  // state is per group.

  __device__ static int glane() { return lane() & 15; }
  __device__ static int group() { return lane() >> 4; }
  // the value v of lane src (ds_bpermute; each lane its own source)
  __device__ static int bperm(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }

  // A reference plane as a motion search reads it: its base at the macroblock's top-left sample
  // and its row stride, taken from the slice parameters once per search (read inside the SAD loop
  // they were loaded from memory again before every SAD).
  struct PlaneAt {
    const uint8_t *mb;  // the macroblock's sample (0, 0) in the padded plane
    int sw;
  };
  __device__ PlaneAt plane_at(int pi) const {
    const int sw = P.mb_w * 16 + 2 * PAD;
    return PlaneAt{P.hp[pi] + (size_t)(mby * 16 + PAD) * sw + (mbx * 16 + PAD), sw};
  }

  // The w x h source block at (bx, by) of the macroblock as the SADs of one search read it, over the
  // group: its lanes take the block's dwords (four samples of a row) in turn, at most four each. Set
  // once per search: each lane's source dwords and the addresses of its dwords in the integer plane
  // at vector (0, 0), so that a candidate adds one offset to them.
  struct SadBlock {
    const uint8_t *row[4];
    uint32_t src[4];
    int sw, nd;
    uint32_t a0;  // the byte alignment of row[] (the same for all: the row stride is a multiple of 4)
  };
  __device__ SadBlock sad_block(const PlaneAt &pa, int bx, int by, int w, int h) const {
    SadBlock sb;
    const int lwd = __builtin_ctz(w) - 2;
    sb.sw = pa.sw;
    sb.nd = (w * h) >> 2;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = glane() + 16 * k, rr = i >> lwd, q = i & ((1 << lwd) - 1);
      sb.row[k] = pa.mb + (ptrdiff_t)(by + rr) * pa.sw + (bx + 4 * q);
      sb.src[k] = i < sb.nd ? *reinterpret_cast<const uint32_t *>(&L.src_y[(by + rr) * 16 + bx + 4 * q]) : 0u;
    }
    sb.a0 = (uint32_t)((uintptr_t)pa.mb & 3);
    return sb;
  }

  // SAD of the block against the integer plane at the integer part of (mx, my) (enc/frame.rs:662-671):
  // all loads issued before the first is used; v_sad_u8 of the source dword and the reference read as
  // two aligned dwords byte-aligned (the plane's border covers the dword past a block's row).
  __device__ int sad_g(const SadBlock &sb, int mx, int my) const {
    typedef IA_GLOBAL_AS const uint32_t GU32;
    const int off = (my >> 2) * sb.sw + (mx >> 2);
    const uint32_t sh = (sb.a0 + (uint32_t)off) & 3;
    const ptrdiff_t offa = off - (int)sh;
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      lo[k] = hi[k] = 0;
      if (glane() + 16 * k < sb.nd) {
        GU32 *d = reinterpret_cast<GU32 *>(as_global(sb.row[k] + offa));
        lo[k] = d[0];
        hi[k] = d[1];
      }
    }
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) v = __builtin_amdgcn_sad_u8(sb.src[k], __builtin_amdgcn_alignbyte(hi[k], lo[k], sh), v);
    return group_sum((int)v, 16);
  }

  // quarter-sample prediction of the w x h block at picture (x, y) from reference r into out
  // (row stride ostride); enc/subpel.rs:122-145
  // (the parameters read once, a lane's at most four sample pairs loaded before any is used)
  __device__ void luma_pred(int r, int x, int y, int w, int h, int mx, int my, uint8_t *out, int ostride) const {
    typedef IA_GLOBAL_AS const uint8_t GU8;
    const int pw = P.mb_w * 16, ph = P.mb_h * 16, sw = pw + 2 * PAD, sh = ph + 2 * PAD;
    const uint8_t *base = P.hp[r];
    const size_t n = (size_t)sw * sh;
    const int fx = mx & 3, fy = my & 3;
    const int ix = clampi(x + (mx >> 2), -w - 3, pw + 1), iy = clampi(y + (my >> 2), -h - 3, ph + 1);
    const uint32_t m = qmix(fy, fx);
    GU8 *a = as_global(base + (m & 3) * n + (size_t)(iy + ((m >> 3) & 1) + PAD) * sw + (ix + ((m >> 2) & 1) + PAD));
    GU8 *b = as_global(base + ((m >> 4) & 3) * n + (size_t)(iy + ((m >> 7) & 1) + PAD) * sw + (ix + ((m >> 6) & 1) + PAD));
    const int lw = __builtin_ctz(w), nn = w * h;
    int va[4], vb[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = lane() + 64 * k, rr = i >> lw, c = i & (w - 1);
      va[k] = vb[k] = 0;
      if (i < nn) {
        va[k] = a[(size_t)rr * sw + c];
        vb[k] = b[(size_t)rr * sw + c];
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = lane() + 64 * k, rr = i >> lw, c = i & (w - 1);
      if (i < nn) out[rr * ostride + c] = (uint8_t)((va[k] + vb[k] + 1) >> 1);
    }
  }

  // SATD of the w x h source block at (bx, by) of plane `pl` (0 luma, 1/2 chroma) against pred
  // (row stride pstride), enc/frame.rs:641-659
  __device__ int satd(int pl, int bx, int by, int w, int h, const uint8_t *pred, int pstride) const {
    const int nbx = w >> 2, nb = nbx * (h >> 2);
    int s = 0;
    if (lane() < nb) {
      const int ox = (lane() % nbx) * 4, oy = (lane() / nbx) * 4;
      const uint8_t *src = pl == 0 ? L.src_y : L.src_c[pl - 1];
      const int ss = pl == 0 ? 16 : 8;
      int d[16];
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++) d[rr * 4 + c] = (int)src[(by + oy + rr) * ss + bx + ox + c] - (int)pred[(oy + rr) * pstride + ox + c];
      s = satd4(d);
    }
    return wave_sum(s);
  }

  // The 4x4 Hadamard transform of the 16 lanes of each 16-lane group (lane 16 q + i holds sample i:
  // row i >> 2, column i & 3), by butterflies between lanes i and i ^ 1, i ^ 2 (quad_perm), i ^ 4
  // (row_half_mirror then quad_perm [3, 2, 1, 0]) and i ^ 8 (row_mirror then row_half_mirror). The
  // lanes end with the transform's outputs in another order than its matrix's, which a sum of
  // magnitudes (the SATD) does not see.
  __device__ static int hadamard16(int v, int i) {
    int p = __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
    v = (i & 1) ? p - v : v + p;
    p = __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);
    v = (i & 2) ? p - v : v + p;
    p = __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false), 0x1b, 0xf, 0xf, false);
    v = (i & 4) ? p - v : v + p;
    p = __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false), 0x141, 0xf, 0xf, false);
    v = (i & 8) ? p - v : v + p;
    return v;
  }

  // v of lane i ^ K within each 16-lane group (K 1, 2, 3 by quad_perm; 4, 8, 12 by row mirrors and
  // quad_perm)
  template <int K>
  __device__ static int lane_xor(int v) {
    if constexpr (K == 1) return __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);
    if constexpr (K == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);
    if constexpr (K == 3) return __builtin_amdgcn_update_dpp(0, v, 0x1b, 0xf, 0xf, false);
    if constexpr (K == 4)
      return __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false), 0x1b, 0xf, 0xf, false);
    if constexpr (K == 8)
      return __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false), 0x141, 0xf, 0xf, false);
    return __builtin_amdgcn_update_dpp(0, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false), 0x1b, 0xf, 0xf, false);
  }
  // One dimension of the 4x4 forward core transform (enc/frame.rs:2443-2459) over the four lanes
  // k = 0..3 that lane_xor<X3> and lane_xor<X1> pair as k ^ 3 and k ^ 1: sums and differences of
  // samples 0, 3 and 1, 2, then of those; position k ends with frequency perm4(k).
  template <int X3, int X1>
  __device__ static int fwd4(int v, int k) {
    int p = lane_xor<X3>(v);
    v = k < 2 ? v + p : p - v;  // a0 + a3, a1 + a2, a1 - a2, a0 - a3
    p = lane_xor<X1>(v);
    return k == 0 ? v + p : k == 1 ? p - v : k == 2 ? p - 2 * v : 2 * v + p;
  }
  // One dimension of the inverse core transform (h264/idct.rs:8-32, wrapping) over four lanes holding
  // frequencies perm4(k) (paired k ^ 1, then k ^ 3): position k ends with sample k.
  template <int X1, int X3>
  __device__ static int inv4(int v, int k) {
    int p = lane_xor<X1>(v);
    v = k == 0 ? wadd(v, p) : k == 1 ? wsub(p, v) : k == 2 ? wsub(p >> 1, v) : wadd(v, p >> 1);  // e0, e1, e2, e3
    p = lane_xor<X3>(v);
    return k < 2 ? wadd(v, p) : wsub(p, v);
  }

  // SATD of a 4x4 block of sample differences (each within +-255): the Hadamard transform with rows 0
  // and 2, and rows 1 and 3, packed in the halves of a dword, so each add or subtract handles two rows
  // (every intermediate is within +-16 * 255, which 16 bits hold). In the column pass both halves
  // of the four sums end as the same column total, and one half is taken.
  typedef int16_t S16x2 __attribute__((vector_size(4)));
  __device__ static S16x2 abs16x2(S16x2 v) {
    const S16x2 m = v >> 15;
    return (v ^ m) - m;
  }
  __device__ static int satd4(const int *d) {
    S16x2 t[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      S16x2 p[4];
#pragma unroll
      for (int c = 0; c < 4; c++) p[c] = S16x2{(int16_t)d[h * 4 + c], (int16_t)d[(h + 2) * 4 + c]};
      const S16x2 s0 = p[0] + p[1], s1 = p[2] + p[3], d0 = p[0] - p[1], d1 = p[2] - p[3];
      t[h][0] = s0 + s1;
      t[h][1] = s0 - s1;
      t[h][2] = d0 - d1;
      t[h][3] = d0 + d1;
    }
    int sum = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const S16x2 sv = t[0][c] + t[1][c], dv = t[0][c] - t[1][c];  // (s0, s1), (d0, d1)
      const S16x2 sx = S16x2{sv[1], sv[0]}, dx = S16x2{dv[1], dv[0]};
      const S16x2 tot = abs16x2(sv + sx) + abs16x2(sv - sx) + abs16x2(dv - dx) + abs16x2(dv + dx);
      sum += tot[0];
    }
    return sum >> 1;
  }

  __device__ int ref_bits(int r) const {
    if (P.nref <= 1) return 0;
    if (P.nref == 2) return 1;
    int v = r + 1, lg = 0;
    while (v >>= 1) lg++;
    return 2 * lg + 1;
  }

  __device__ int col_mv(int16_t *mv) const {
    if (!P.rmbs0) return 0;
    const ia_mbinfo *col = &P.rmbs0[mby * P.mb_w + mbx];
    if (col->kind != K_INTER) return 0;
    mv[0] = col->mv[5][0];
    mv[1] = col->mv[5][1];
    return 1;
  }

  __device__ static int fs_index(int x, int y, int w, int h) {
    if (w == 4 && h == 4) return 0;
    if (w == 4 && h == 2) return 1 + y / 2;
    if (w == 2 && h == 4) return 3 + x / 2;
    int b8 = (y / 2) * 2 + x / 2;
    if (w == 2 && h == 2) return 5 + b8;
    if (w == 2 && h == 1) return 9 + 2 * b8 + (y & 1);
    if (w == 1 && h == 2) return 17 + 2 * b8 + (x & 1);
    return 25 + 4 * b8 + (y & 1) * 2 + (x & 1);
  }

  // The candidates of one sub-sample pass: around its centre the eight half-sample neighbours (pass 0,
  // the centre itself first) or the four quarter-sample ones (pass 1), each clamped to the search
  // window. Candidate k's offset from the centre plus one, per coordinate, in two bits of a word.
  struct SubPass {
    int cx0, cy0, pass, lo_x, hi_x, lo_y, hi_y;
    static constexpr int OFF[2][9][2] = {{{0, 0}, {-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}},
                                         {{0, -1}, {-1, 0}, {1, 0}, {0, 1}}};
    static constexpr uint32_t pack(int ps, int c) {
      uint32_t v = 0;
      for (int k = 0; k < 9; k++) v |= (uint32_t)(OFF[ps][k][c] + 1) << (2 * k);
      return v;
    }
    __device__ int n() const { return pass == 0 ? 9 : 4; }
    __device__ void at(int k, int &x, int &y) const {
      constexpr uint32_t X0 = pack(0, 0), Y0 = pack(0, 1), X1 = pack(1, 0), Y1 = pack(1, 1);
      const uint32_t fx = pass == 0 ? X0 : X1, fy = pass == 0 ? Y0 : Y1;
      const int step = pass == 0 ? 2 : 1;
      x = clampi(cx0 + ((int)((fx >> (2 * k)) & 3) - 1) * step, lo_x, hi_x);
      y = clampi(cy0 + ((int)((fy >> (2 * k)) & 3) - 1) * step, lo_y, hi_y);
    }
  };


  // The costs SATD + lambda * mvd_bits that the sub-sample refinement gives n <= 9 candidate vectors
  // of the w x h partition at (bxp, byp) of the macroblock (picture position (x0, y0)) against
  // reference r, over the group: a lane takes one 4x4 block of one candidate's quarter-sample
  // prediction (enc/subpel.rs:122-145, as luma_pred forms it), 16 / blocks candidates at a time, and
  // each candidate's blocks are summed within its lanes.
  __device__ void cost_sub_g(int r, int x0, int y0, int bxp, int byp, int w, int h, const SubPass &sp,
                             const int16_t *mvp, int *costs) const {
    typedef IA_GLOBAL_AS const uint8_t GU8;
    typedef IA_GLOBAL_AS const uint32_t GU32u __attribute__((aligned(1)));  // a dword at any byte address
    const int pw = P.mb_w * 16, ph = P.mb_h * 16, sw = pw + 2 * PAD, sh = ph + 2 * PAD;
    const int plane = sw * sh;  // the four planes of a reference are within 2^31 bytes
    const int nbx = w >> 2, nb = nbx * (h >> 2), lnb = __builtin_ctz(nb), per = 16 >> lnb;
    const int lam = P.lambda, gb = lane() & ~15, n = sp.n();
    // the lane's block, the same for every candidate: its sample (0, 0) in the planes at vector (0, 0)
    const int blk = glane() & (nb - 1), ox = (blk % nbx) * 4, oy = (blk / nbx) * 4;
    const uint8_t *lb = P.hp[r] + (ptrdiff_t)(oy + PAD) * sw + (ox + PAD);
    const uint8_t *ls = &L.src_y[(byp + oy) * 16 + bxp + ox];
    for (int base = 0; base < n; base += per) {
      const int gi = base + (glane() >> lnb);
      int sv = 0;
      if (gi < n) {
        int vx, vy;
        sp.at(gi, vx, vy);
        const int fx = vx & 3, fy = vy & 3;
        const int ix = clampi(x0 + (vx >> 2), -w - 3, pw + 1), iy = clampi(y0 + (vy >> 2), -h - 3, ph + 1);
        const uint32_t m = qmix(fy, fx);
        GU8 *a = as_global(lb + ((int)(m & 3) * plane + (iy + (int)((m >> 3) & 1)) * sw + ix + (int)((m >> 2) & 1)));
        GU8 *b = as_global(lb + ((int)((m >> 4) & 3) * plane + (iy + (int)((m >> 7) & 1)) * sw + ix + (int)((m >> 6) & 1)));
        // the rounded mean of the two planes' samples four at a time: (a + b + 1) >> 1 of each byte is
        // (a | b) - ((a ^ b) >> 1), with no carry between bytes
        int d[16];
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
          const uint32_t pa = *reinterpret_cast<GU32u *>(a + (size_t)rr * sw);
          const uint32_t pb = *reinterpret_cast<GU32u *>(b + (size_t)rr * sw);
          const uint32_t pv = (pa | pb) - (((pa ^ pb) >> 1) & 0x7f7f7f7fu);
          const uint32_t sr = *reinterpret_cast<const uint32_t *>(ls + rr * 16);
#pragma unroll
          for (int c = 0; c < 4; c++) d[rr * 4 + c] = (int)((sr >> (8 * c)) & 255) - (int)((pv >> (8 * c)) & 255);
        }
        sv = satd4(d);
      }
      sv = group_sum(sv, nb);
      for (int k = 0; k < per && base + k < n; k++) {
        int cx, cy;
        sp.at(base + k, cx, cy);
        costs[base + k] = bperm(sv, gb + (k << lnb)) + lam * mvd_bits(cx, cy, mvp[0], mvp[1]);
      }
    }
  }

  // ---- motion search (enc/frame.rs:949-1019) ----
  //
  // The four groups run the same code: their control flow is kept uniform (each loop runs while any
  // group still needs it, and a group that is done only stops taking the results), so every DPP
  // reduction is reached by all lanes. The geometry of a search is the same in all groups; the
  // reference, predictors and candidates are the group's.

  // any lane's predicate holds (the same in every lane)
  __device__ static bool any_lane(bool p) { return __ballot(p) != 0; }

  // list l (B slices); r the reference index in it; pi the half-sample planes searched
  __device__ int motion_search(int r, int bx, int by, int bw, int bh, const int16_t *mvp, const int16_t (*extra)[2],
                               int nextra, int16_t *out, int l = 0) {
    const int lam = P.lambda, pi = P.b_slice ? l : r;
    const int x0 = mbx * 16 + bx * 4, y0 = mby * 16 + by * 4, w = bw * 4, h = bh * 4;
    const int lo_x = max(-(x0 + 24) * 4, -2048), hi_x = min((P.mb_w * 16 - x0 + 8) * 4, 2047);
    const int lo_y = max(-(y0 + 24) * 4, -512), hi_y = min((P.mb_h * 16 - y0 + 8) * 4, 511);
    // the candidate list across the group's lanes: lane j holds candidate j (at most 16), its two
    // components as the halves of a dword
    int lst = 0, nc = 0;
    auto push = [&](int vx, int vy) {
      const int cx = (clampi(vx, lo_x, hi_x) + 2) & ~3, cy = (clampi(vy, lo_y, hi_y) + 2) & ~3;
      if (glane() == nc) lst = (int)(((uint32_t)cx & 0xffffu) | ((uint32_t)cy << 16));
      nc++;
    };
    push(mvp[0], mvp[1]);
    push(0, 0);  // (the window holds the zero vector, which the rounding keeps)
    {
      // the neighbours' vectors only: positions outside the macroblock never read the partition state
      const int NB[3][2] = {{-1, 0}, {0, -1}, {4, -1}};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        int rr;
        int16_t mv[2];
        if (nb_motion(nullptr, nullptr, 0, NB[k][0], NB[k][1], &rr, mv, l) && rr == r) push(mv[0], mv[1]);
      }
    }
    for (int k = 0; k < nextra; k++) push(extra[k][0], extra[k][1]);
    if (P.fs && !P.b_slice) {
      const ia_fsresult *f = &P.fs[((size_t)r * P.mb_w * P.mb_h + (size_t)mby * P.mb_w + mbx) * IA_NPART + fs_index(bx, by, bw, bh)];
      push(f->mv[0], f->mv[1]);
    }
    const int bxp = bx * 4, byp = by * 4, gb = lane() & ~15;
    const SadBlock sb = sad_block(plane_at(pi), bxp, byp, w, h);
    int bestx = 0, besty = 0, best_cost = 0x7fffffff;
    for (int k = 0; any_lane(k < nc); k++) {
      const bool on = k < nc;
      const int pk = bperm(lst, gb + (k & 15));
      const int cx = on ? (int)(int16_t)(pk & 0xffff) : 0, cy = on ? (pk >> 16) : 0;
      const int cost = sad_g(sb, cx, cy) + lam * mvd_bits(cx, cy, mvp[0], mvp[1]);
      if (on && cost < best_cost) { bestx = cx; besty = cy; best_cost = cost; }
    }
    // the realtime decisions skip the 16-sample step (enc/frame.rs:973)
    for (int si = P.fast ? 1 : 0; si < 3; si++) {
      const int step = 16 >> si, iters = step == 4 ? 32 : 4;
      bool going = true;  // this group's search at this step has not yet stopped
      for (int it = 0; it < iters && any_lane(going); it++) {
        int moved = 0;
        for (int d = 0; d < 4; d++) {
          int dx = d == 0 ? -step : d == 1 ? step : 0, dy = d == 2 ? -step : d == 3 ? step : 0;
          int cx = clampi(bestx + dx, lo_x, hi_x), cy = clampi(besty + dy, lo_y, hi_y);
          int cost = sad_g(sb, cx, cy) + lam * mvd_bits(cx, cy, mvp[0], mvp[1]);
          if (going && cost < best_cost) { bestx = cx; besty = cy; best_cost = cost; moved = 1; }
        }
        if (!moved) going = false;
      }
    }
    // realtime: stop at an integer vector that already matches (enc/frame.rs:990-992)
    const bool stop = P.fast && best_cost < w * h / 8 + lam;
    // sub-sample refinement by SATD: 8 half samples around, then 4 quarter samples (realtime coding
    // keeps to the half samples, enc/frame.rs:1006). Each pass probes points around a centre fixed for
    // the pass, so its candidates' costs are computed together and then taken in the reference's order
    // with its strict minimum; the first pass carries the centre's own cost.
    for (int pass = 0; pass < (P.fast ? 1 : 2) && any_lane(!stop); pass++) {
      const SubPass sp{bestx, besty, pass, lo_x, hi_x, lo_y, hi_y};
      const int n = sp.n();
      int cost[9];
      cost_sub_g(pi, x0, y0, bxp, byp, w, h, sp, mvp, cost);
      if (!stop) {
        int k0 = 0;
        if (pass == 0) {
          best_cost = cost[0];
          k0 = 1;
        }
        for (int k = k0; k < n; k++) {
          int cx, cy;
          sp.at(k, cx, cy);
          if (cost[k] < best_cost) { bestx = cx; besty = cy; best_cost = cost[k]; }
        }
      }
    }
    out[0] = (int16_t)bestx;
    out[1] = (int16_t)besty;
    return best_cost - lam * mvd_bits(bestx, besty, mvp[0], mvp[1]);
  }

  // enc/frame.rs:911-940; the search state lives in the group's L.ms.st_*, and each sub-partitioning
  // is tried on a copy of it in L.ms.tr_*, lane i of the group keeping 4x4 block i (raster): its entry
  // of the state is copied, painted when a sub-partition covering it is found, and taken as the kept
  // state's when the sub-partitioning is the best so far.
  __device__ int sub_search(int r, int k, const int16_t *extra, uint32_t *done, int *cost_out) {
    const int lam = P.lambda, ox = (k & 1) * 2, oy = (k >> 1) * 2, g = group(), i = glane(), ix = i & 3, iy = i >> 2;
    int best_s = -1, best_cost = 0;
    int16_t bm0 = L.ms.st_mvs[g][i][0], bm1 = L.ms.st_mvs[g][i][1];
    int8_t br = L.ms.st_refs[g][i];
    uint32_t bdone = 0;
    for (int s = 0; s < (P.sub8 ? 4 : 1); s++) {
      L.ms.tr_mvs[g][i][0] = L.ms.st_mvs[g][i][0];
      L.ms.tr_mvs[g][i][1] = L.ms.st_mvs[g][i][1];
      L.ms.tr_refs[g][i] = L.ms.st_refs[g][i];
      sync();
      uint32_t td = *done;
      int cost = lam * (s == 0 ? 0 : 2);
      for (int q = 0; q < NSUBS[s]; q++) {
        const int x = ox + SUBS[s][q][0], y = oy + SUBS[s][q][1], w = SUBS[s][q][2], h = SUBS[s][q][3];
        int16_t mvp[2], m[2];
        pred_mv(L.ms.tr_mvs[g], L.ms.tr_refs[g], td, x, y, w, 0, r, mvp);
        int cc = motion_search(r, x, y, w, h, mvp, (const int16_t(*)[2])extra, 1, m);
        cost += cc + lam * mvd_bits(m[0], m[1], mvp[0], mvp[1]);
        if (ix >= x && ix < x + w && iy >= y && iy < y + h) {
          L.ms.tr_mvs[g][i][0] = m[0];
          L.ms.tr_mvs[g][i][1] = m[1];
          L.ms.tr_refs[g][i] = (int8_t)r;
        }
        for (int yy = y; yy < y + h; yy++) td |= ((1u << w) - 1) << (yy * 4 + x);
        sync();
      }
      if (best_s < 0 || cost < best_cost) {
        best_s = s;
        best_cost = cost;
        bdone = td;
        bm0 = L.ms.tr_mvs[g][i][0];
        bm1 = L.ms.tr_mvs[g][i][1];
        br = L.ms.tr_refs[g][i];
      }
      sync();
    }
    // the kept sub-partitioning's state becomes the group's
    L.ms.st_mvs[g][i][0] = bm0;
    L.ms.st_mvs[g][i][1] = bm1;
    L.ms.st_refs[g][i] = br;
    sync();
    *done = bdone;
    *cost_out = best_cost;
    return best_s;
  }

  struct Inter {
    int part;
    int16_t mvs[16][2];
    uint8_t refs[4], sub[4];
    int cost;
  };

  // enc/frame.rs:852-906, for the group's reference r
  __device__ void best_inter_ref(int r, Inter &best) {
    const int lam = P.lambda, rb = lam * ref_bits(r), g = group();
    uint8_t refs[4] = {(uint8_t)r, (uint8_t)r, (uint8_t)r, (uint8_t)r};
    int16_t mvp[2], mv[2], col[1][2];
    mv_pred(r, mvp);
    int ncol = col_mv(col[0]);
    int cost = motion_search(r, 0, 0, 4, 4, mvp, (const int16_t(*)[2])col, ncol, mv);
    best.part = 0;
    for (int i = 0; i < 16; i++) { best.mvs[i][0] = mv[0]; best.mvs[i][1] = mv[1]; }
    for (int k = 0; k < 4; k++) { best.refs[k] = (uint8_t)r; best.sub[k] = 0; }
    best.cost = cost + lam * (mvd_bits(mv[0], mv[1], mvp[0], mvp[1]) + 1) + rb;
    if (P.fast) return;  // realtime: 16x16 only (enc/frame.rs:861-863)
    if (glane() == 0)
      for (int i = 0; i < 16; i++) { L.ms.st_mvs[g][i][0] = 0; L.ms.st_mvs[g][i][1] = 0; L.ms.st_refs[g][i] = 0; }
    sync();
    uint32_t done = 0;
    uint8_t sub[4] = {0, 0, 0, 0};
    int cost8 = lam * 9;
    bool alive = true;  // this group's 8x8 total is still below its best
    for (int k = 0; k < 4 && any_lane(alive); k++) {
      int cc;
      const int sk = sub_search(r, k, mv, &done, &cc);
      if (alive) {
        sub[k] = (uint8_t)sk;
        cost8 += cc + rb;
        if (cost8 >= best.cost) alive = false;
      }
    }
    const bool use8 = alive && cost8 < best.cost;
    if (!any_lane(use8)) return;
    if (use8) {
      Rect rs[16];
      int n = rects(3, sub, rs);
      best.part = 3;
      for (int i = 0; i < 16; i++) { best.mvs[i][0] = 0; best.mvs[i][1] = 0; }
      for (int k = 0; k < n; k++) {
        best.mvs[k][0] = L.ms.st_mvs[g][rs[k].y * 4 + rs[k].x][0];
        best.mvs[k][1] = L.ms.st_mvs[g][rs[k].y * 4 + rs[k].x][1];
      }
      for (int k = 0; k < 4; k++) best.sub[k] = sub[k];
      best.cost = cost8;
    }
    int16_t at8[4][2];
    for (int b = 0; b < 4; b++) {
      at8[b][0] = L.ms.st_mvs[g][(b / 2) * 8 + (b % 2) * 2][0];
      at8[b][1] = L.ms.st_mvs[g][(b / 2) * 8 + (b % 2) * 2][1];
    }
    for (int part = 1; part < 3; part++) {
      int16_t mvs[16][2] = {};
      int pc = lam * 3;
      bool palive = use8;
      const uint8_t z4[4] = {0, 0, 0, 0};
      for (int k = 0; k < 2 && any_lane(palive); k++) {
        int16_t ps[16][2] = {}, m[2], cands[2][2];
        part_mvps(part, z4, mvs, refs, ps);
        int b0 = part == 1 ? 2 * k : k, b1 = part == 1 ? 2 * k + 1 : k + 2;
        cands[0][0] = at8[b0][0]; cands[0][1] = at8[b0][1];
        cands[1][0] = at8[b1][0]; cands[1][1] = at8[b1][1];
        const uint8_t *pr = PARTS[part][k];
        int cc = motion_search(r, pr[0], pr[1], pr[2], pr[3], ps[k], (const int16_t(*)[2])cands, 2, m);
        if (palive) {
          mvs[k][0] = m[0];
          mvs[k][1] = m[1];
          pc += cc + lam * mvd_bits(m[0], m[1], ps[k][0], ps[k][1]) + rb;
          if (pc >= best.cost) palive = false;
        }
      }
      if (palive && pc < best.cost) {
        best.part = part;
        for (int i = 0; i < 16; i++) { best.mvs[i][0] = mvs[i][0]; best.mvs[i][1] = mvs[i][1]; }
        for (int k = 0; k < 4; k++) best.sub[k] = 0;
        best.cost = pc;
      }
    }
    for (int k = 0; k < 4; k++) best.refs[k] = (uint8_t)r;
  }

  // the references four at a time, reference c0 + g in group g (a group beyond the last reference
  // repeats the chunk's first and is not taken); their results through LDS, then taken in reference
  // order, the first least cost (enc/frame.rs:840-848)
  __device__ void best_inter(Inter &out) {
    const int g = group();
    for (int c0 = 0; c0 < P.nref; c0 += 4) {
      const int r = c0 + g < P.nref ? c0 + g : c0;
      Inter t;
      best_inter_ref(r, t);
      if (glane() == 0) {
        SearchRes &q = L.ms.res[g];
        q.cost = t.cost;
        q.part = t.part;
        for (int i = 0; i < 16; i++) { q.mvs[i][0] = t.mvs[i][0]; q.mvs[i][1] = t.mvs[i][1]; }
        for (int k = 0; k < 4; k++) { q.refs[k] = t.refs[k]; q.sub[k] = t.sub[k]; }
      }
      sync();
      for (int k = 0; k < 4 && c0 + k < P.nref; k++) {
        const SearchRes &q = L.ms.res[k];
        if (c0 + k == 0 || q.cost < out.cost) {
          out.cost = q.cost;
          out.part = q.part;
          for (int i = 0; i < 16; i++) { out.mvs[i][0] = q.mvs[i][0]; out.mvs[i][1] = q.mvs[i][1]; }
          for (int j = 0; j < 4; j++) { out.refs[j] = q.refs[j]; out.sub[j] = q.sub[j]; }
        }
      }
      sync();
    }
  }

  // ---- intra 16x16 choice (enc/frame.rs:792-834) ----

  __device__ void i16_pred(int mode, uint8_t *out) {
    const int ta = up(), la = mbx > 0;
    // lanes over samples; edges from the canvas
    for (int i = lane(); i < 256; i += 64) {
      const int x = i & 15, y = i >> 4;
      int v;
      if (mode == 0) v = cyv(x, -1);
      else if (mode == 1) v = cyv(-1, y);
      else if (mode == 2) {
        int st = 0, sl = 0;
        if (ta) for (int k = 0; k < 16; k++) st += cyv(k, -1);
        if (la) for (int k = 0; k < 16; k++) sl += cyv(-1, k);
        if (ta && la) v = (st + sl + 16) >> 5;
        else if (la) v = (sl + 8) >> 4;
        else if (ta) v = (st + 8) >> 4;
        else v = 128;
      } else {
        int h = 0, vv = 0;
        for (int k = 0; k < 8; k++) {
          h += (k + 1) * ((int)cyv(8 + k, -1) - (int)cyv(6 - k, -1));
          vv += (k + 1) * ((int)cyv(-1, 8 + k) - (int)cyv(-1, 6 - k));
        }
        const int a = 16 * ((int)cyv(-1, 15) + (int)cyv(15, -1)), b = (5 * h + 32) >> 6, c = (5 * vv + 32) >> 6;
        v = clampi((a + b * (x - 7) + c * (y - 7) + 16) >> 5, 0, 255);
      }
      out[i] = (uint8_t)v;
    }
  }

  __device__ int best_i16(int *cost_out) {
    const int ta = up(), la = mbx > 0;
    int best_m = 2, best_c = 0x7fffffff;
    for (int mode = 0; mode < 4; mode++) {
      int ok = mode == 0 ? ta : mode == 1 ? la : mode == 2 ? 1 : (ta && la);
      if (!ok) continue;
      i16_pred(mode, L.pred);
      sync();
      int cost = satd(0, 0, 0, 16, 16, L.pred, 16) + P.lambda * (mode == 2 ? 0 : 1);
      sync();
      if (cost < best_c) { best_m = mode; best_c = cost; }
    }
    *cost_out = best_c;
    return best_m;
  }

  // ---- transforms (enc/frame.rs:113-141, h264/idct.rs) ----

  __device__ static void fdct4(const int *d, int *o) {
    int t[16];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int *a = d + r * 4;
      int s03 = a[0] + a[3], d03 = a[0] - a[3], s12 = a[1] + a[2], d12 = a[1] - a[2];
      t[r * 4] = s03 + s12;
      t[r * 4 + 1] = 2 * d03 + d12;
      t[r * 4 + 2] = s03 - s12;
      t[r * 4 + 3] = d03 - 2 * d12;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      int s03 = t[c] + t[12 + c], d03 = t[c] - t[12 + c], s12 = t[4 + c] + t[8 + c], d12 = t[4 + c] - t[8 + c];
      o[c] = s03 + s12;
      o[4 + c] = 2 * d03 + d12;
      o[8 + c] = s03 - s12;
      o[12 + c] = d03 - 2 * d12;
    }
  }

  __device__ static int quant(int c, int mf, int f, int qbits) {
    long long a = ((long long)(c < 0 ? -(long long)c : c) * mf + f) >> qbits;
    return c < 0 ? -(int)a : (int)a;
  }

  __device__ static int class4(int i) {
    int a = i & 1, b = (i >> 2) & 1;
    return (a == 0 && b == 0) ? 0 : (a == 1 && b == 1) ? 1 : 2;
  }

  __device__ static int class8(int p) {
    int i = p / 8, j = p % 8;
    if (i % 4 == 0 && j % 4 == 0) return 0;
    if (i % 2 == 1 && j % 2 == 1) return 1;
    if (i % 4 == 2 && j % 4 == 2) return 2;
    if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) return 3;
    if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) return 4;
    return 5;
  }

  // enc/frame.rs:1025-1046: levels in scan order (lv), dequantised raster coefficients (co), DC
  __device__ static int quant_block(const int *d, int qp, int intra, int skip_dc, int16_t *lv, int32_t *co) {
    int c[16];
    fdct4(d, c);
    const int qbits = 15 + qp / 6, f = (1 << qbits) / (intra ? 3 : 6), m = qp % 6;
    for (int k = 0; k < 16; k++) {
      lv[k] = 0;
      co[k] = 0;
    }
    for (int k = 0; k < 16; k++) {
      const int p = ZIG4[k];
      if (skip_dc && p == 0) continue;
      const int l = quant(c[p], IA_MF[m][class4(p)], f, qbits);
      lv[k] = (int16_t)l;
      co[p] = shl(l * IA_V[m][class4(p)], qp / 6);
    }
    return c[0];
  }

  __device__ static void fdct8_1d(const int *d, int *o) {
    int s07 = d[0] + d[7], s16 = d[1] + d[6], s25 = d[2] + d[5], s34 = d[3] + d[4];
    int a0 = s07 + s34, a1 = s16 + s25, a2 = s07 - s34, a3 = s16 - s25;
    int d07 = d[0] - d[7], d16 = d[1] - d[6], d25 = d[2] - d[5], d34 = d[3] - d[4];
    int a4 = d16 + d25 + (d07 + (d07 >> 1));
    int a5 = d07 - d34 - (d25 + (d25 >> 1));
    int a6 = d07 + d34 - (d16 + (d16 >> 1));
    int a7 = d16 - d25 + (d34 + (d34 >> 1));
    o[0] = a0 + a1;
    o[1] = a4 + (a7 >> 2);
    o[2] = a2 + (a3 >> 1);
    o[3] = a5 + (a6 >> 2);
    o[4] = a0 - a1;
    o[5] = a6 - (a5 >> 2);
    o[6] = (a2 >> 1) - a3;
    o[7] = (a4 >> 2) - a7;
  }

  __device__ static int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
  __device__ static int wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }

  // h264/idct.rs:8-32 into 4x4 samples at dst (stride)
  __device__ static void idct4_add(const int32_t *c, uint8_t *dst, int stride) {
    int t[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int32_t *d = c + i * 4;
      int e0 = wadd(d[0], d[2]), e1 = wsub(d[0], d[2]), e2 = wsub(d[1] >> 1, d[3]), e3 = wadd(d[1], d[3] >> 1);
      t[i * 4] = wadd(e0, e3);
      t[i * 4 + 1] = wadd(e1, e2);
      t[i * 4 + 2] = wsub(e1, e2);
      t[i * 4 + 3] = wsub(e0, e3);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int g0 = wadd(t[j], t[8 + j]), g1 = wsub(t[j], t[8 + j]), g2 = wsub(t[4 + j] >> 1, t[12 + j]), g3 = wadd(t[4 + j], t[12 + j] >> 1);
      int hh[4] = {wadd(g0, g3), wadd(g1, g2), wsub(g1, g2), wsub(g0, g3)};
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint8_t *p = dst + i * stride + j;
        *p = (uint8_t)clampi(*p + (wadd(hh[i], 32) >> 6), 0, 255);
      }
    }
  }

  __device__ static void row8(const int *d, int *o) {
    int a0 = wadd(d[0], d[4]), a4 = wsub(d[0], d[4]), a2 = wsub(d[2] >> 1, d[6]), a6 = wadd(d[2], d[6] >> 1);
    int b0 = wadd(a0, a6), b2 = wadd(a4, a2), b4 = wsub(a4, a2), b6 = wsub(a0, a6);
    int a1 = wsub(wsub(wsub(d[5], d[3]), d[7]), d[7] >> 1);
    int a3 = wsub(wsub(wadd(d[1], d[7]), d[3]), d[3] >> 1);
    int a5 = wadd(wadd(wsub(d[7], d[1]), d[5]), d[5] >> 1);
    int a7 = wadd(wadd(wadd(d[3], d[5]), d[1]), d[1] >> 1);
    int b1 = wadd(a1, a7 >> 2), b7 = wsub(a7, a1 >> 2), b3 = wadd(a3, a5 >> 2), b5 = wsub(a3 >> 2, a5);
    o[0] = wadd(b0, b7);
    o[1] = wadd(b2, b5);
    o[2] = wadd(b4, b3);
    o[3] = wadd(b6, b1);
    o[4] = wsub(b6, b1);
    o[5] = wsub(b4, b3);
    o[6] = wsub(b2, b5);
    o[7] = wsub(b0, b7);
  }

  __device__ static void hadamard4(const int32_t *c, int32_t *o) {
    int t[16];
    for (int i = 0; i < 4; i++) {
      const int32_t *d = c + i * 4;
      int s0 = wadd(d[0], d[1]), s1 = wsub(d[0], d[1]), s2 = wadd(d[2], d[3]), s3 = wsub(d[2], d[3]);
      t[i * 4] = wadd(s0, s2);
      t[i * 4 + 1] = wsub(s0, s2);
      t[i * 4 + 2] = wsub(s1, s3);
      t[i * 4 + 3] = wadd(s1, s3);
    }
    for (int j = 0; j < 4; j++) {
      int s0 = wadd(t[j], t[4 + j]), s1 = wsub(t[j], t[4 + j]), s2 = wadd(t[8 + j], t[12 + j]), s3 = wsub(t[8 + j], t[12 + j]);
      o[j] = wadd(s0, s2);
      o[4 + j] = wsub(s0, s2);
      o[8 + j] = wsub(s1, s3);
      o[12 + j] = wadd(s1, s3);
    }
  }

  __device__ static int decimate_score(const int16_t *lv) {
    const int RUN[16] = {3, 2, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int score = 0, i = 15;
    while (i >= 0 && lv[i] == 0) i--;
    while (i >= 0) {
      if (iabs(lv[i]) > 1) return 9;
      i--;
      int run = 0;
      while (i >= 0 && lv[i] == 0) { i--; run++; }
      score += RUN[min(run, 15)];
    }
    return score;
  }

  __device__ static int decimate_score8(const int16_t *lv) {
    const int RUN[25] = {3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0};
    int score = 0, i = 63;
    while (i >= 0 && lv[i] == 0) i--;
    while (i >= 0) {
      if (iabs(lv[i]) > 1) return 9;
      i--;
      int run = 0;
      while (i >= 0 && lv[i] == 0) { i--; run++; }
      score += RUN[min(run, 24)];
    }
    return score;
  }

  __device__ static int count_nz(const int16_t *v, int n) {
    int k = 0;
    for (int i = 0; i < n; i++) k += v[i] != 0;
    return k;
  }

  // ---- 8x8 transforms split over the lanes (the arithmetic of quant8 / idct8_add, enc/frame.rs:88-105,
  // h264/idct.rs:55-74) ----

  // Blocks b0..b0+nb-1 of L.co8 hold raster residuals: forward transform (a lane per row, then per
  // column), then quantisation and dequantisation (a lane per coefficient). Levels land in lv[b] in
  // zig-zag order, the dequantised coefficients back in L.co8[b] (raster); nz[b - b0] the nonzero levels.
  __device__ void quant8_par(int b0, int nb, int qp, int intra, int16_t (*lv)[64], int *nz) {
    const int l = lane();
    if (l < nb * 8) {
      int32_t *row = L.co8[b0 + (l >> 3)] + (l & 7) * 8;
      int t[8], o[8];
      for (int k = 0; k < 8; k++) t[k] = row[k];
      fdct8_1d(t, o);
      for (int k = 0; k < 8; k++) row[k] = o[k];
    }
    sync();
    if (l < nb * 8) {
      int32_t *col = L.co8[b0 + (l >> 3)] + (l & 7);
      int t[8], o[8];
      for (int k = 0; k < 8; k++) t[k] = col[k * 8];
      fdct8_1d(t, o);
      for (int k = 0; k < 8; k++) col[k * 8] = o[k];
    }
    sync();
    const int m = qp % 6, qbits = 16 + qp / 6, f = (1 << qbits) / (intra ? 3 : 6);
    const int pz = ZIG8[l], cls = class8(pz), mf = IA_MF8[m][cls], dq = 16 * IA_DEQUANT8_COEFF_INIT[m][cls];
    for (int b = b0; b < b0 + nb; b++) {
      const int lvl = quant(L.co8[b][pz], mf, f, qbits);
      lv[b][l] = (int16_t)lvl;
      int co = 0;
      if (lvl != 0) {
        const int v = lvl * dq;
        co = qp >= 36 ? shl(v, qp / 6 - 6) : (v + (1 << (5 - qp / 6))) >> (6 - qp / 6);
      }
      L.co8[b][pz] = co;  // this lane alone reads and writes coefficient pz
      nz[b - b0] = count_lanes(lvl != 0);
    }
    sync();
  }

  // the inverse transform of L.co8[b] added to the canvas at (x, y) for the blocks whose bit is set in
  // `mask` (b = 0..3 at (b & 1) * 8, (b >> 1) * 8): a lane per row, then per column
  __device__ void idct8_add_par(unsigned mask) {
    const int l = lane(), b = l >> 3;
    const bool on = l < 32 && ((mask >> b) & 1);
    if (on) {
      int32_t *row = L.co8[b] + (l & 7) * 8;
      int t[8], o[8];
      for (int k = 0; k < 8; k++) t[k] = row[k];
      row8(t, o);
      for (int k = 0; k < 8; k++) row[k] = o[k];
    }
    sync();
    if (on) {
      const int j = l & 7;
      const int32_t *col = L.co8[b] + j;
      int d[8], hh[8];
      for (int i = 0; i < 8; i++) d[i] = col[i * 8];
      row8(d, hh);
      uint8_t *dst = &cyv((b & 1) * 8 + j, (b >> 1) * 8);
      for (int i = 0; i < 8; i++) {
        uint8_t *q = dst + i * 28;
        *q = (uint8_t)clampi(*q + (wadd(hh[i], 32) >> 6), 0, 255);
      }
    }
    sync();
  }

  // ---- coding of the candidates (enc/frame.rs:1050-1918) ----

  __device__ void code_new(int kind) {
    // every lane clears its share of the code
    uint32_t *w = reinterpret_cast<uint32_t *>(&L.code);
    for (int i = lane(); i < (int)(sizeof(ia_mbcode) / 4); i += 64) w[i] = 0;
    sync();
    if (lane() == 0) {
      L.code.kind = (uint8_t)kind;
      for (int i = 0; i < 16; i++) L.code.modes[i] = 2;
    }
    sync();
  }

  __device__ static void block_pos(int n, int *bx, int *by) {
    *bx = ((n >> 2) & 1) * 2 + (n & 1);
    *by = ((n >> 3) & 1) * 2 + ((n >> 1) & 1);
  }

  __device__ static int zorder(int bx, int by) { return ((by >> 1) << 3) | ((bx >> 1) << 2) | ((by & 1) << 1) | (bx & 1); }

  // directional 4x4/8x8 prediction of sample (x, y) (h264/intra.rs:19-127); t(i)/l(j) read the edges
  __device__ static int nxn_sample(int n, int mode, int x, int y, const int *top, const int *left, int tl, int have_top,
                                   int have_left) {
#define T(i) ((i) < 0 ? tl : top[(i)])
#define LF(j) ((j) < 0 ? tl : left[(j)])
    const int shift = n == 4 ? 2 : 3;
    switch (mode) {
      case 0: return T(x);
      case 1: return LF(y);
      case 2: {
        int st = 0, sl = 0;
        for (int i = 0; i < n; i++) { st += T(i); sl += LF(i); }
        if (have_top && have_left) return (st + sl + n) >> (shift + 1);
        if (have_left) return (sl + n / 2) >> shift;
        if (have_top) return (st + n / 2) >> shift;
        return 128;
      }
      case 3:
        if (x == n - 1 && y == n - 1) return (T(2 * n - 2) + 3 * T(2 * n - 1) + 2) >> 2;
        return (T(x + y) + 2 * T(x + y + 1) + T(x + y + 2) + 2) >> 2;
      case 4:
        if (x > y) return (T(x - y - 2) + 2 * T(x - y - 1) + T(x - y) + 2) >> 2;
        if (x < y) return (LF(y - x - 2) + 2 * LF(y - x - 1) + LF(y - x) + 2) >> 2;
        return (T(0) + 2 * tl + LF(0) + 2) >> 2;
      case 5: {
        int z = 2 * x - y;
        if (z >= 0 && z % 2 == 0) return (T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 1) >> 1;
        if (z >= 0) return (T(x - (y >> 1) - 2) + 2 * T(x - (y >> 1) - 1) + T(x - (y >> 1)) + 2) >> 2;
        if (z == -1) return (LF(0) + 2 * tl + T(0) + 2) >> 2;
        return (LF(y - 2 * x - 1) + 2 * LF(y - 2 * x - 2) + LF(y - 2 * x - 3) + 2) >> 2;
      }
      case 6: {
        int z = 2 * y - x;
        if (z >= 0 && z % 2 == 0) return (LF(y - (x >> 1) - 1) + LF(y - (x >> 1)) + 1) >> 1;
        if (z >= 0) return (LF(y - (x >> 1) - 2) + 2 * LF(y - (x >> 1) - 1) + LF(y - (x >> 1)) + 2) >> 2;
        if (z == -1) return (LF(0) + 2 * tl + T(0) + 2) >> 2;
        return (T(x - 2 * y - 1) + 2 * T(x - 2 * y - 2) + T(x - 2 * y - 3) + 2) >> 2;
      }
      case 7:
        if (y % 2 == 0) return (T(x + (y >> 1)) + T(x + (y >> 1) + 1) + 1) >> 1;
        return (T(x + (y >> 1)) + 2 * T(x + (y >> 1) + 1) + T(x + (y >> 1) + 2) + 2) >> 2;
      default: {
        int z = x + 2 * y, edge = 2 * n - 3;
        if (z > edge) return LF(n - 1);
        if (z == edge) return (LF(n - 2) + 3 * LF(n - 1) + 2) >> 2;
        if (z % 2 == 0) return (LF(y + (x >> 1)) + LF(y + (x >> 1) + 1) + 1) >> 1;
        return (LF(y + (x >> 1)) + 2 * LF(y + (x >> 1) + 1) + LF(y + (x >> 1) + 2) + 2) >> 2;
      }
    }
#undef T
#undef LF
  }

  // Intra4x4 sample (x, y) of mode m (not DC, h264/intra.rs:19-127) as taps on the edge samples E (E[0..3]
  // the left column from the bottom up, E[4] the corner, E[5..12] the row above and above-right): the
  // sample is (w0 E[i0] + w1 E[i1] + w2 E[i2] + 2) >> 2, packed i0 | i1 << 4 | i2 << 8 | w0 << 12 |
  // w1 << 15 | w2 << 18. Each mode's cases of nxn_sample in that form, so that lanes of different
  // modes run one branch-free evaluation.
  __device__ static uint32_t tap4(int m, int x, int y) {
    int i0 = 0, i1 = 0, i2 = 0, w0 = 0, w1 = 0, w2 = 0;
    auto t3 = [&](int a) {  // E[a] + 2 E[a + 1] + E[a + 2]
      i0 = a; i1 = a + 1; i2 = a + 2; w0 = 1; w1 = 2; w2 = 1;
    };
    auto t2 = [&](int a, int b) {  // (E[a] + E[b] + 1) >> 1
      i0 = a; i1 = b; w0 = 2; w1 = 2;
    };
    switch (m) {
      case 0: i0 = 5 + x; w0 = 4; break;
      case 1: i0 = 3 - y; w0 = 4; break;
      case 3:
        if (x == 3 && y == 3) { i0 = 11; i1 = 12; w0 = 1; w1 = 3; }
        else t3(5 + x + y);
        break;
      case 4: t3(3 + x - y); break;
      case 5: {
        const int z = 2 * x - y, k = x - (y >> 1);
        if (z >= 0 && !(z & 1)) t2(4 + k, 5 + k);
        else if (z >= 0) t3(3 + k);
        else if (z == -1) t3(3);
        else t3(4 + z);
        break;
      }
      case 6: {
        const int z = 2 * y - x, k = y - (x >> 1);
        if (z >= 0 && !(z & 1)) t2(4 - k, 3 - k);
        else if (z >= 0) t3(3 - k);
        else if (z == -1) t3(3);
        else t3(2 - z);
        break;
      }
      case 7: {
        const int k = x + (y >> 1);
        if (!(y & 1)) t2(5 + k, 6 + k);
        else t3(5 + k);
        break;
      }
      case 8: {
        const int z = x + 2 * y, k = y + (x >> 1);
        if (z > 5) { i0 = 0; w0 = 4; }
        else if (z == 5) { i0 = 1; i1 = 0; w0 = 1; w1 = 3; }
        else if (!(z & 1)) t2(3 - k, 2 - k);
        else t3(1 - k);
        break;
      }
      default: break;
    }
    return (uint32_t)(i0 | i1 << 4 | i2 << 8 | w0 << 12 | w1 << 15 | w2 << 18);
  }

  __device__ static int mode_ok(int mode, int have_top, int have_left, int have_tl) {
    if (mode == 0 || mode == 3 || mode == 7) return have_top;
    if (mode == 1 || mode == 8) return have_left;
    if (mode == 2) return 1;
    return have_top && have_left && have_tl;
  }

  // enc/frame.rs:1050-1138: Intra4x4 into L.code and the canvas; returns the SATD cost
  __device__ int code_i4() {
    code_new(K_I4);
    const int ta = up(), la = mbx > 0, tla = mbx > 0 && ta, tra = ta && mbx + 1 < P.mb_w;
    const int qp = P.qp;
    const int l = lane();
    int cost = 0;
    // the lane's taps for its mode of each pass below (mode 4 pass + l / 16, sample l % 16)
    uint32_t taps[3];
#pragma unroll
    for (int pass = 0; pass < 3; pass++) taps[pass] = tap4(4 * pass + (l >> 4), l & 3, (l >> 2) & 3);
    // the coefficient the lane holds after the forward transform, its scan position and its
    // quantiser and dequantiser (enc/frame.rs:126-138)
    const int coef = perm4((l >> 2) & 3) * 4 + perm4(l & 3);
    const int scan = (int)((IZIG4 >> (4 * coef)) & 15);
    const int qbits = 15 + qp / 6, qf = (1 << qbits) / 3;
    const int mf = IA_MF[qp % 6][class4(coef)], vq = IA_V[qp % 6][class4(coef)];
    for (int n = 0; n < 16; n++) {
      int bx, by;
      block_pos(n, &bx, &by);
      const int x = bx * 4, y = by * 4;
      const int have_left = bx > 0 || la, have_top = by > 0 || ta;
      int have_tl, have_tr;
      if (bx > 0 && by > 0) have_tl = 1;
      else if (by > 0) have_tl = la;
      else if (bx > 0) have_tl = ta;
      else have_tl = tla;
      if (by == 0) have_tr = bx < 3 ? ta : tra;
      else if (bx == 3) have_tr = 0;
      else have_tr = zorder(bx + 1, by - 1) < zorder(bx, by);
      // the edge samples E[j] (tap4) read from the canvas as the taps need them: E[0..3] the column
      // to the left from the bottom up, E[4] the corner, E[5..12] the row above and above-right
      // (sample 3 repeated when the above-right is missing). A mode whose edges are missing is not
      // allowed, so the taps of an allowed mode with a nonzero weight read only present edges.
      auto edge = [&](int j) -> int {
        const int ex = j < 5 ? x - 1 : x + (j - 5 < 4 || have_tr ? j - 5 : 3), ey = j < 4 ? y + 3 - j : y - 1;
        return cyv(ex, ey);
      };
      auto tap_at = [&](uint32_t t) {
        return ((int)((t >> 12) & 7) * edge(t & 15) + (int)((t >> 15) & 7) * edge((t >> 4) & 15) +
                (int)((t >> 18) & 7) * edge((t >> 8) & 15) + 2) >> 2;
      };
      const int predicted = pred_i4_mode(L.code.modes, bx, by);
      // DC (mode 2) for every lane at once (nxn_sample's case 2)
      int dc;
      {
        int st = 0, sl = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          st += cyv(x + i, y - 1);
          sl += cyv(x - 1, y + i);
        }
        dc = have_top && have_left ? (st + sl + 4) >> 3 : have_left ? (sl + 2) >> 2 : have_top ? (st + 2) >> 2 : 128;
      }
      // the nine modes' SATDs, four modes a pass: lane 16 g + i takes sample i (row i >> 2, column
      // i & 3) of mode 4 pass + g; the Hadamard across the group's lanes (hadamard16), the absolute
      // values summed over the 16 lanes (enc/frame.rs:641-659)
      int cpass[3];
#pragma unroll
      for (int pass = 0; pass < 3; pass++) {
        const int m = 4 * pass + (l >> 4), i = l & 15, xx = i & 3, yy = i >> 2;
        const bool ok = m < 9 && mode_ok(m, have_top, have_left, have_tl);
        int v = 0;
        if (ok) v = (int)L.src_y[(y + yy) * 16 + x + xx] - (m == 2 ? dc : tap_at(taps[pass]));
        const int sum = group_sum(iabs(hadamard16(v, i)), 16) >> 1;
        cpass[pass] = ok ? sum + P.lambda * (m == predicted ? 1 : 4) : 0x7fffffff;
      }
      // the first mode of least cost (strict <, in mode order)
      int best_m = 2, best_c = 0x7fffffff;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const int ck = __builtin_amdgcn_readlane(cpass[k >> 2], 16 * (k & 3));
        if (ck < best_c) { best_c = ck; best_m = k; }
      }
      cost += best_c;
      // the block coded with it (enc/frame.rs:1025-1046, h264/idct.rs:8-32) in the registers of the
      // 16-lane groups: lane i holds sample (r, c) = (i >> 2, i & 3) and after the forward transform
      // coefficient (perm4(r), perm4(c)); lanes i < 16 quantise it and write the block
      {
        const int i = l & 15, r = i >> 2, c = i & 3;
        const int pv = best_m == 2 ? dc : tap_at(tap4(best_m, c, r));
        int v = (int)L.src_y[(y + r) * 16 + x + c] - pv;
        v = fwd4<3, 1>(v, c);   // along the row
        v = fwd4<12, 4>(v, r);  // along the column
        const int lvl = quant(v, mf, qf, qbits);
        if (l < 16) L.code.luma[by * 4 + bx][scan] = (int16_t)lvl;
        const int nz = count_lanes(l < 16 && lvl != 0);
        int rec = pv;
        if (nz) {
          int d = shl(lvl * vq, qp / 6);
          d = inv4<1, 3>(d, c);   // the rows first, as h264/idct.rs
          d = inv4<4, 12>(d, r);
          rec = clampi(pv + (wadd(d, 32) >> 6), 0, 255);
        }
        if (l < 16) cyv(x + c, y + r) = (uint8_t)rec;
        if (l == 0) {
          L.code.nnz[by * 4 + bx] = (uint8_t)nz;
          L.code.modes[by * 4 + bx] = (uint8_t)best_m;
        }
      }
      sync();
    }
    if (lane() == 0) {
      int cbp = 0;
      for (int i8 = 0; i8 < 4; i8++) {
        int bx = (i8 & 1) * 2, by = (i8 >> 1) * 2, any = 0;
        for (int dy = 0; dy < 2; dy++)
          for (int dx = 0; dx < 2; dx++) any |= L.code.nnz[(by + dy) * 4 + bx + dx] > 0;
        if (any) cbp |= 1 << i8;
      }
      L.code.cbp = (uint8_t)cbp;
    }
    sync();
    return cost;
  }

  // enc/frame.rs:1140-1197
  __device__ void code_i16(int mode) {
    code_new(K_I16);
    if (lane() == 0) L.code.i16_mode = (uint8_t)mode;
    i16_pred(mode, L.pred);
    sync();
    const int qp = P.qp;
    // lane b < 16: block b's AC levels and DC
    if (lane() < 16) {
      const int b = lane(), bx = b % 4, by = b / 4;
      int d[16];
      for (int i = 0; i < 16; i++) {
        const int o = (by * 4 + (i >> 2)) * 16 + bx * 4 + (i & 3);
        d[i] = (int)L.src_y[o] - (int)L.pred[o];
      }
      L.dcs[b] = quant_block(d, qp, 1, 1, L.code.luma[b], L.co[b]);
    }
    sync();
    if (lane() == 0) {
      int32_t h[16], dl[16] = {};
      hadamard4(L.dcs, h);
      const int qbits = 16 + qp / 6, f = (1 << qbits) / 3, m = qp % 6;
      for (int k = 0; k < 16; k++) {
        const int p = ZIG4[k];
        const int l = quant(h[p] >> 1, IA_MF[m][0], f, qbits);
        L.code.dc[k] = (int16_t)l;
        dl[p] = l;
      }
      int ac = 0;
      for (int b = 0; b < 16; b++) ac |= count_nz(L.code.luma[b], 16) > 0;
      if (!ac)
        for (int b = 0; b < 16; b++)
          for (int k = 0; k < 16; k++) { L.code.luma[b][k] = 0; L.co[b][k] = 0; }
      L.code.cbp = ac ? 15 : 0;
      int32_t ff[16];
      hadamard4(dl, ff);
      const int lsd = 16 * IA_V[m][0];
      for (int b = 0; b < 16; b++)
        L.co[b][0] = qp >= 36 ? shl(ff[b] * lsd, qp / 6 - 6) : (ff[b] * lsd + (1 << (5 - qp / 6))) >> (6 - qp / 6);
    }
    sync();
    for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
    sync();
    if (lane() < 16) {
      const int b = lane(), bx = b % 4, by = b / 4;
      int any = 0;
      for (int k = 0; k < 16; k++) any |= L.co[b][k] != 0;
      if (any) idct4_add(L.co[b], &cyv(bx * 4, by * 4), 28);
      L.code.nnz[b] = (uint8_t)count_nz(L.code.luma[b], 16);
    }
    sync();
  }

  // enc/frame.rs:1856-1918: chroma of the candidate against L.cpred
  __device__ void code_chroma(int intra) {
    const int qp = P.qpc, m = qp % 6;
    // lane (c, blk) < 8: the 4x4 block's AC levels and DC
    if (lane() < 8) {
      const int ci = lane() >> 2, blk = lane() & 3, bx = blk & 1, by = blk >> 1;
      int d[16];
      for (int i = 0; i < 16; i++) {
        const int o = (by * 4 + (i >> 2)) * 8 + bx * 4 + (i & 3);
        d[i] = (int)L.src_c[ci][o] - (int)L.cpred[ci][o];
      }
      L.dcs[lane()] = quant_block(d, qp, intra, 1, L.code.cac[ci][blk], L.co[lane()]);
      L.score[lane()] = decimate_score(L.code.cac[ci][blk]);
    }
    sync();
    if (lane() == 0) {
      int any_dc = 0, any_ac = 0;
      for (int ci = 0; ci < 2; ci++) {
        const int *dcs = L.dcs + ci * 4;
        const int f[4] = {dcs[0] + dcs[1] + dcs[2] + dcs[3], dcs[0] - dcs[1] + dcs[2] - dcs[3], dcs[0] + dcs[1] - dcs[2] - dcs[3],
                          dcs[0] - dcs[1] - dcs[2] + dcs[3]};
        const int qbits = 16 + qp / 6, fr = (1 << qbits) / (intra ? 3 : 6);
        for (int k = 0; k < 4; k++) L.code.cdc[ci][k] = (int16_t)quant(f[k], IA_MF[m][0], fr, qbits);
        any_dc |= count_nz(L.code.cdc[ci], 4) > 0;
        for (int blk = 0; blk < 4; blk++) any_ac |= count_nz(L.code.cac[ci][blk], 16) > 0;
      }
      int ac_score = 0;
      for (int k = 0; k < 8; k++) ac_score += L.score[k];
      if (!intra && ac_score < 7) any_ac = 0;
      if (!any_ac)
        for (int k = 0; k < 8; k++)
          for (int q = 0; q < 16; q++) { L.code.cac[k >> 2][k & 3][q] = 0; L.co[k][q] = 0; }
      const int cbpc = any_ac ? 2 : (any_dc ? 1 : 0);
      L.code.cbp |= (uint8_t)(cbpc << 4);
      const int lsd = 16 * IA_V[m][0];
      for (int ci = 0; ci < 2; ci++) {
        const int16_t *l = L.code.cdc[ci];
        const int f[4] = {l[0] + l[1] + l[2] + l[3], l[0] - l[1] + l[2] - l[3], l[0] + l[1] - l[2] - l[3], l[0] - l[1] - l[2] + l[3]};
        for (int blk = 0; blk < 4; blk++) L.co[ci * 4 + blk][0] = cbpc > 0 ? shl(f[blk] * lsd, qp / 6) >> 5 : 0;
      }
    }
    sync();
    for (int i = lane(); i < 128; i += 64) {
      const int ci = i >> 6, o = i & 63;
      ccv(ci, o & 7, o >> 3) = L.cpred[ci][o];
    }
    sync();
    if (lane() < 8) {
      const int ci = lane() >> 2, blk = lane() & 3, bx = blk & 1, by = blk >> 1;
      L.code.cnnz[ci][blk] = (uint8_t)count_nz(L.code.cac[ci][blk], 16);
      int any = 0;
      for (int k = 0; k < 16; k++) any |= L.co[lane()][k] != 0;
      if (any) idct4_add(L.co[lane()], &ccv(ci, bx * 4, by * 4), 12);
    }
    sync();
  }

  // enc/frame.rs:1254-1304 (the luma against L.pred with the 8x8 transform)
  __device__ void code_luma8() {
    const int qp = P.qp;
    // residuals of the four 8x8 blocks (block b = i >> 6, raster sample i & 63)
    for (int i = lane(); i < 256; i += 64) {
      const int b = i >> 6, k = i & 63, o = ((b >> 1) * 8 + (k >> 3)) * 16 + (b & 1) * 8 + (k & 7);
      L.co8[b][k] = (int)L.src_y[o] - (int)L.pred[o];
    }
    sync();
    int nz[4];
    quant8_par(0, 4, qp, 0, L.code.luma8, nz);
    if (lane() < 4) L.score[lane()] = decimate_score8(L.code.luma8[lane()]);
    for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
    sync();
    const int total = L.score[0] + L.score[1] + L.score[2] + L.score[3];
    unsigned coded = 0;
    int n[4];
    for (int b = 0; b < 4; b++) {
      const bool kept = !(L.score[b] < 4 || total < 6);
      n[b] = kept ? nz[b] : 0;
      if (!kept)
        for (int k = lane(); k < 64; k += 64) L.code.luma8[b][k] = 0;
      if (n[b] > 0) coded |= 1u << b;
    }
    sync();
    if (coded) idct8_add_par(coded);
    if (lane() < 4) {
      const int i8 = lane();
      if (!(L.score[i8] < 4 || total < 6))
        for (int j = 0; j < 4; j++) L.code.nnz[((i8 >> 1) * 2 + (j >> 1)) * 4 + (i8 & 1) * 2 + (j & 1)] = (uint8_t)n[i8];
      L.ival[i8] = n[i8];
    }
    sync();
    if (lane() == 0) {
      int cbp = L.code.cbp;
      for (int i8 = 0; i8 < 4; i8++)
        if (!(L.score[i8] < 4 || total < 6) && L.ival[i8] > 0) cbp |= 1 << i8;
      L.code.cbp = (uint8_t)cbp;
      L.code.t8x8 = (cbp & 15) != 0;
    }
    sync();
  }

  // enc/frame.rs:1783-1805: luma and chroma prediction of an inter macroblock into L.pred/L.cpred
  __device__ void inter_pred(int part, const uint8_t *sub, const int16_t (*mvs)[2], const uint8_t *refs) {
    Rect rs[16];
    const int n = rects(part, sub, rs);
    const int cw = P.mb_w * 8, ch = P.mb_h * 8;
    for (int k = 0; k < n; k++) {
      const int bx = rs[k].x, by = rs[k].y, w = rs[k].w, h = rs[k].h;
      const int r = refs[ref_of(part, k, bx, by)];
      const int mx = mvs[k][0], my = mvs[k][1];
      luma_pred(r, mbx * 16 + bx * 4, mby * 16 + by * 4, w * 4, h * 4, mx, my, L.pred + by * 4 * 16 + bx * 4, 16);
      // chroma: eighth-sample bilinear from a window clamped to the plane (h264/inter.rs:377-422)
      const int fx = mx & 7, fy = my & 7;
      const int x0 = mbx * 8 + bx * 2 + (mx >> 3), y0 = mby * 8 + by * 2 + (my >> 3);
      const int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
      const int cwn = w * 2, chn = h * 2;
      for (int i = lane(); i < 2 * cwn * chn; i += 64) {
        const int ci = i / (cwn * chn), j = i % (cwn * chn), rr = j / cwn, c = j % cwn;
        const uint8_t *pl = ci == 0 ? P.rcb[r] : P.rcr[r];
        const int X = x0 + c, Y = y0 + rr;
#define S(xx, yy) ((int)pl[(size_t)clampi((yy), 0, ch - 1) * cw + clampi((xx), 0, cw - 1)])
        const int v = (wa * S(X, Y) + wb * S(X + 1, Y) + wc * S(X, Y + 1) + wd * S(X + 1, Y + 1) + 32) >> 6;
#undef S
        L.cpred[ci][(by * 2 + rr) * 8 + bx * 2 + c] = (uint8_t)v;
      }
    }
    sync();
  }

  // enc/frame.rs:1217-1252 (residual of an inter macroblock against L.pred/L.cpred)
  __device__ void code_inter_residual(int t8) {
    if (t8) {
      code_luma8();
    } else {
      const int qp = P.qp;
      if (lane() < 16) {
        const int b = lane(), bx = b % 4, by = b / 4;
        int d[16];
        for (int i = 0; i < 16; i++) {
          const int o = (by * 4 + (i >> 2)) * 16 + bx * 4 + (i & 3);
          d[i] = (int)L.src_y[o] - (int)L.pred[o];
        }
        quant_block(d, qp, 0, 0, L.code.luma[b], L.co[b]);
        L.score[b] = decimate_score(L.code.luma[b]);
      }
      sync();
      for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
      sync();
      if (lane() < 16) {
        const int b = lane(), bx = b % 4, by = b / 4, i8 = (by / 2) * 2 + bx / 2;
        int s8[4] = {0, 0, 0, 0};
        for (int q = 0; q < 16; q++) s8[((q / 4) / 2) * 2 + (q % 4) / 2] += L.score[q];
        const int total = s8[0] + s8[1] + s8[2] + s8[3];
        if (s8[i8] < 4 || total < 6)
          for (int k = 0; k < 16; k++) { L.code.luma[b][k] = 0; L.co[b][k] = 0; }
        const int nz = count_nz(L.code.luma[b], 16);
        L.code.nnz[b] = (uint8_t)nz;
        if (nz > 0) idct4_add(L.co[b], &cyv(bx * 4, by * 4), 28);
        L.ival[b] = nz;
      }
      sync();
      if (lane() == 0) {
        int cbp = 0;
        for (int b = 0; b < 16; b++)
          if (L.ival[b] > 0) cbp |= 1 << (((b / 4) / 2) * 2 + (b % 4) / 2);
        L.code.cbp = (uint8_t)cbp;
      }
      sync();
    }
    code_chroma(0);
  }

  // enc/frame.rs:1201-1214
  __device__ void code_inter(int part, const int16_t (*mvs)[2], const uint8_t *refs, const uint8_t *sub, int t8) {
    code_new(K_INTER);
    int16_t mvps[16][2] = {};
    part_mvps(part, sub, mvs, refs, mvps);
    if (lane() == 0) {
      L.code.part = (uint8_t)part;
      for (int i = 0; i < 16; i++) {
        L.code.mvs[i][0] = mvs[i][0];
        L.code.mvs[i][1] = mvs[i][1];
        L.code.mvds[i][0] = (int16_t)(mvs[i][0] - mvps[i][0]);
        L.code.mvds[i][1] = (int16_t)(mvs[i][1] - mvps[i][1]);
      }
      for (int k = 0; k < 4; k++) { L.code.refs[k] = refs[k]; L.code.sub[k] = sub[k]; }
    }
    sync();
    inter_pred(part, sub, mvs, refs);
    code_inter_residual(t8);
  }

  // enc/frame.rs:1307-1388
  __device__ void code_i8() {
    code_new(K_I8);
    if (lane() == 0) L.code.t8x8 = 1;
    sync();
    const int ta = up(), la = mbx > 0, tla = mbx > 0 && ta, tra = ta && mbx + 1 < P.mb_w;
    const int qp = P.qp;
    for (int i8 = 0; i8 < 4; i8++) {
      const int bx = i8 & 1, by = i8 >> 1, x = bx * 8, y = by * 8;
      const int have_left = bx > 0 || la, have_top = by > 0 || ta;
      int have_tl;
      if (bx > 0 && by > 0) have_tl = 1;
      else if (by > 0) have_tl = la;
      else if (bx > 0) have_tl = ta;
      else have_tl = tla;
      const int have_tr = i8 == 0 ? ta : i8 == 1 ? tra : i8 == 2 ? 1 : 0;
      // the edges, smoothed (h264/intra.rs:130-159), into LDS: lane i < 16 the row above (8-15 above
      // right, or sample 7 repeated), 16-23 the column to the left, 24 the corner; zero where missing
      {
        const int l = lane();
        auto pt = [&](int i) { return have_top ? (int)cyv(x + (i < 8 || have_tr ? i : 7), y - 1) : 0; };
        auto pl = [&](int j) { return have_left ? (int)cyv(x - 1, y + j) : 0; };
        const int ptl = have_tl ? (int)cyv(x - 1, y - 1) : 0;
        if (l < 16) {
          int v = 0;
          if (have_top) {
            if (l == 0) v = have_tl ? (ptl + 2 * pt(0) + pt(1) + 2) >> 2 : (3 * pt(0) + pt(1) + 2) >> 2;
            else if (l == 15) v = (pt(14) + 3 * pt(15) + 2) >> 2;
            else v = (pt(l - 1) + 2 * pt(l) + pt(l + 1) + 2) >> 2;
          }
          L.etop[l] = v;
        } else if (l < 24) {
          const int j = l - 16;
          int v = 0;
          if (have_left) {
            if (j == 0) v = have_tl ? (ptl + 2 * pl(0) + pl(1) + 2) >> 2 : (3 * pl(0) + pl(1) + 2) >> 2;
            else if (j == 7) v = (pl(6) + 3 * pl(7) + 2) >> 2;
            else v = (pl(j - 1) + 2 * pl(j) + pl(j + 1) + 2) >> 2;
          }
          L.eleft[j] = v;
        } else if (l == 24) {
          int v = ptl;
          if (have_tl) {
            if (have_top && have_left) v = (pt(0) + 2 * ptl + pl(0) + 2) >> 2;
            else if (have_top) v = (3 * ptl + pt(0) + 2) >> 2;
            else if (have_left) v = (3 * ptl + pl(0) + 2) >> 2;
          }
          L.etl = v;
        }
      }
      sync();
      const int *top = L.etop, *left = L.eleft, tl = L.etl;
      const int predicted = pred_i4_mode(L.code.modes, bx * 2, by * 2);
      // the nine modes' SATDs, a mode a pass (no divergence): lane 16 q + i takes sample i of the
      // 4x4 quadrant q; the Hadamard across the quadrant's 16 lanes (hadamard16), each quadrant's
      // SATD summed over its 16 lanes, the mode's over the quadrants
      int best_m = 2, best_c = 0x7fffffff;
      {
        const int l = lane(), qd = l >> 4, i = l & 15, xx = (qd & 1) * 4 + (i & 3), yy = (qd >> 1) * 4 + (i >> 2);
        const int sv = L.src_y[(y + yy) * 16 + x + xx];
        for (int m = 0; m < 9; m++) {
          if (!mode_ok(m, have_top, have_left, have_tl)) continue;
          const int v = sv - nxn_sample(8, m, xx, yy, top, left, tl, have_top, have_left);
          const int q = group_sum(iabs(hadamard16(v, i)), 16) >> 1;
          const int sat = __builtin_amdgcn_readlane(q, 0) + __builtin_amdgcn_readlane(q, 16) + __builtin_amdgcn_readlane(q, 32) +
                          __builtin_amdgcn_readlane(q, 48);
          const int c = sat + P.lambda * (m == predicted ? 1 : 4);
          if (c < best_c) { best_c = c; best_m = m; }
        }
      }
      // the prediction into the canvas and the residual, a lane per sample
      {
        const int i = lane();
        const int pv = nxn_sample(8, best_m, i & 7, i >> 3, top, left, tl, have_top, have_left);
        L.co8[i8][i] = (int)L.src_y[(y + (i >> 3)) * 16 + x + (i & 7)] - pv;
        cyv(x + (i & 7), y + (i >> 3)) = (uint8_t)pv;
      }
      sync();
      int n;
      quant8_par(i8, 1, qp, 1, L.code.luma8, &n);
      if (n > 0) idct8_add_par(1u << i8);
      if (lane() == 0) {
        if (n > 0) L.code.cbp |= (uint8_t)(1 << i8);
        for (int j = 0; j < 4; j++) {
          const int b = (by * 2 + (j >> 1)) * 4 + bx * 2 + (j & 1);
          L.code.modes[b] = (uint8_t)best_m;
          L.code.nnz[b] = (uint8_t)n;
        }
      }
      sync();
    }
  }

  // enc/frame.rs:1807-1852
  __device__ void code_intra_chroma() {
    const int la = mbx > 0, ta = up();
    // each mode's SATD over both components, a (mode, component) a pass over the 64 samples (no
    // divergence): lane 16 q + i takes sample i of 4x4 block q; Hadamard as for the luma blocks
    int best_m = 0, best_c = 0x7fffffff;
    {
      const int l = lane(), qd = l >> 4, i = l & 15, xx = (qd & 1) * 4 + (i & 3), yy = (qd >> 1) * 4 + (i >> 2);
      for (int mm = 0; mm < 4; mm++) {
        const int ok = mm == 0 ? 1 : mm == 1 ? la : mm == 2 ? ta : (la && ta);
        if (!ok) continue;
        int cost = P.lambda * (mm == 0 ? 1 : 3);
        for (int ci = 0; ci < 2; ci++) {
          const int v = (int)L.src_c[ci][yy * 8 + xx] - chroma_sample(ci, mm, xx, yy, ta, la);
          const int q = group_sum(iabs(hadamard16(v, i)), 16) >> 1;
          cost += __builtin_amdgcn_readlane(q, 0) + __builtin_amdgcn_readlane(q, 16) + __builtin_amdgcn_readlane(q, 32) +
                  __builtin_amdgcn_readlane(q, 48);
        }
        if (cost < best_c) { best_c = cost; best_m = mm; }
      }
    }
    for (int i = lane(); i < 128; i += 64) {
      const int c2 = i >> 6, o = i & 63;
      L.cpred[c2][o] = (uint8_t)chroma_sample(c2, best_m, o & 7, o >> 3, ta, la);
    }
    if (lane() == 0) L.code.chroma_mode = (uint8_t)best_m;
    sync();
    code_chroma(1);
  }

  // chroma intra prediction of sample (x, y) from the canvas (h264/intra.rs:212-287, 8x8)
  __device__ int chroma_sample(int ci, int mode, int x, int y, int ta, int la) {
    if (mode == 1) return la ? ccv(ci, -1, y) : 0;
    if (mode == 2) return ta ? ccv(ci, x, -1) : 0;
    if (mode == 0) {
      const int bx = x >> 2, by = y >> 2;
      int st = 0, sl = 0;
      for (int i = 0; i < 4; i++) {
        st += ta ? ccv(ci, bx * 4 + i, -1) : 0;
        sl += la ? ccv(ci, -1, by * 4 + i) : 0;
      }
      if ((bx == 0 && by == 0) || (bx > 0 && by > 0)) {
        if (ta && la) return (st + sl + 4) >> 3;
        if (la) return (sl + 2) >> 2;
        if (ta) return (st + 2) >> 2;
        return 128;
      }
      if (bx > 0) {
        if (ta) return (st + 2) >> 2;
        if (la) return (sl + 2) >> 2;
        return 128;
      }
      if (la) return (sl + 2) >> 2;
      if (ta) return (st + 2) >> 2;
      return 128;
    }
    // plane
    const int tl = (la && ta) ? ccv(ci, -1, -1) : 0;
    auto tt = [&](int i) { return i < 0 ? tl : (ta ? (int)ccv(ci, i, -1) : 0); };
    auto lf = [&](int j) { return j < 0 ? tl : (la ? (int)ccv(ci, -1, j) : 0); };
    int hh = 0, vv = 0;
    for (int k = 0; k < 4; k++) {
      hh += (k + 1) * (tt(4 + k) - tt(2 - k));
      vv += (k + 1) * (lf(4 + k) - lf(2 - k));
    }
    const int a = 16 * (lf(7) + tt(7)), b = (34 * hh + 32) >> 6, c = (34 * vv + 32) >> 6;
    return clampi((a + b * (x - 3) + c * (y - 3) + 16) >> 5, 0, 255);
  }

  // ---- rate-distortion decision (enc/frame.rs:1391-1513) ----

  __device__ void save_rec(uint8_t *rec) {
    for (int i = lane(); i < 384; i += 64) {
      if (i < 256) rec[i] = cyv(i & 15, i >> 4);
      else {
        const int ci = (i - 256) >> 6, o = (i - 256) & 63;
        rec[i] = ccv(ci, o & 7, o >> 3);
      }
    }
  }

  __device__ void restore_rec(const uint8_t *rec) {
    for (int i = lane(); i < 384; i += 64) {
      if (i < 256) cyv(i & 15, i >> 4) = rec[i];
      else {
        const int ci = (i - 256) >> 6, o = (i - 256) & 63;
        ccv(ci, o & 7, o >> 3) = rec[i];
      }
    }
  }

  __device__ long long ssd_mb() {
    int s = 0;
    for (int i = lane(); i < 384; i += 64) {
      int a, b;
      if (i < 256) { a = L.src_y[i]; b = cyv(i & 15, i >> 4); }
      else {
        const int ci = (i - 256) >> 6, o = (i - 256) & 63;
        a = L.src_c[ci][o];
        b = ccv(ci, o & 7, o >> 3);
      }
      s += (a - b) * (a - b);
    }
    return (long long)wave_sum(s);  // at most 384 * 255^2 < 2^31
  }

  // ---- CABAC bins (enc/frame.rs:2028-2279, enc/cabac.rs:150-246), counted over the lanes ----
  //
  // A macroblock's decision bins fall into chains that share no context: the header syntax
  // (mb_skip_flag, mb_type, sub_mb_type, ref_idx, coded_block_pattern, transform_size_8x8_flag,
  // intra_chroma_pred_mode, mb_qp_delta), the Intra4x4/8x8 mode flags, each component of the motion
  // vector differences, and per residual block category its coded_block_flags, each significance
  // and last position (for 8x8 blocks each significance context and each last context) and each
  // level context. A lane takes one chain and codes its decisions in the writer's order, so every
  // context passes through the states the writer's does and the costs summed over the lanes are the
  // writer's. Bypass and terminate bins cost the same wherever they are counted.

  // per block category: coefficients per block, blocks per macroblock, significance/last lanes
  // (positions 0..max-2; for 8x8 blocks 15 significance and 9 last contexts), level contexts
  __device__ static int cat_max(int cat) { return cat == 0 || cat == 2 ? 16 : cat == 3 ? 4 : cat == 5 ? 64 : 15; }
  __device__ static int cat_blocks(int cat) { return cat == 0 ? 1 : cat == 3 ? 2 : cat == 4 ? 8 : cat == 5 ? 4 : 16; }
  __device__ static int cat_sig_jobs(int cat) { return cat == 5 ? 24 : cat_max(cat) - 1; }
  __device__ static int cat_jobs(int cat) { return (cat != 5) + cat_sig_jobs(cat) + (cat == 3 ? 9 : 10); }

  // one decision on a state held in a register: its cost in 1/256 bits, the next state
  __device__ void dec(uint32_t &s, int bin, uint32_t &bits) const {
    const uint32_t e = L.cab[s * 2 + (bin != (int)(s & 1))];
    bits += e >> 8;
    s = e & 0xffu;
  }
  __device__ void dec_at(uint8_t *st, int ctx, int bin, uint32_t &bits) const {
    uint32_t s = st[ctx];
    dec(s, bin, bits);
    st[ctx] = (uint8_t)s;
  }
  // bins of the k-th order Exp-Golomb code of v (enc/cabac.rs:221-246): a unary prefix of t ones,
  // its terminating zero and k + t suffix bits, t = floor(log2(v / 2^k + 1))
  __device__ static int eg_bins(uint32_t v, int k) {
    const int t = 31 - __clz((int)((v >> k) + 1));
    return 2 * t + 1 + k;
  }

  // Masks (bit = scan position) of the residual blocks that are coded: L.nzm[0][w] their nonzero
  // levels, L.nzm[1][w] the levels above 1 in magnitude. Words 0-3 hold the luma 4x4 blocks 4w..4w+3
  // in coding order (16 bits each), 4-7 the luma 8x8 blocks, 8-9 the chroma AC blocks (component,
  // block) four to a word, 10 the luma DC block (bits 0-15) and the chroma DC blocks (bits 16-19,
  // 20-23); 11-15 stay zero. A lane per coefficient, a ballot per word.
  __device__ void nz_masks() {
    const ia_mbcode &c = L.code;
    const int l = lane(), cbpl = c.cbp & 15, cbpc = c.cbp >> 4, i16 = c.kind == K_I16, t8 = c.t8x8;
    auto put = [&](int w, int v) {
      const uint64_t nz = __ballot(v != 0), g1 = __ballot(v > 1 || v < -1);
      if (l == 0) {
        L.nzm[0][w] = nz;
        L.nzm[1][w] = g1;
      }
    };
    for (int w = 0; w < 4; w++) {
      int v = 0;
      if (!t8 && cbpl) {
        const int n = 4 * w + (l >> 4), pos = l & 15;
        int bx, by;
        block_pos(n, &bx, &by);
        const int16_t *blk = c.luma[by * 4 + bx];
        if ((cbpl >> (n >> 2)) & 1) v = i16 ? (pos < 15 ? blk[pos + 1] : 0) : blk[pos];
      }
      put(w, v);
    }
    for (int b = 0; b < 4; b++) put(4 + b, t8 && ((cbpl >> b) & 1) ? c.luma8[b][l] : 0);
    for (int w = 0; w < 2; w++) {
      const int pos = l & 15;
      put(8 + w, cbpc == 2 && pos < 15 ? c.cac[w][l >> 4][pos + 1] : 0);
    }
    {
      int v = 0;
      if (l < 16) v = i16 ? c.dc[l] : 0;
      else if (l < 24) v = cbpc > 0 ? c.cdc[(l - 16) >> 2][(l - 16) & 3] : 0;
      put(10, v);
    }
    if (l < 10) L.nzm[l / 5][11 + l % 5] = 0;
  }

  // The mask words of a category (four from its first: luma 4x4 0, 8x8 4, chroma AC 8, DC 10) and the
  // mask of its block b from them.
  __device__ static int cat_word(int cat) { return cat == 0 || cat == 3 ? 10 : cat == 4 ? 8 : cat == 5 ? 4 : 0; }
  struct Words {
    uint64_t w0, w1, w2, w3;
    __device__ uint64_t word(int i) const { return i == 0 ? w0 : i == 1 ? w1 : i == 2 ? w2 : w3; }
    __device__ uint64_t block(int cat, int b) const {
      switch (cat) {
        case 0: return w0 & 0xffffu;
        case 3: return (w0 >> (16 + 4 * b)) & 15u;
        case 5: return word(b);
        default: return (word(b >> 2) >> (16 * (b & 3))) & 0xffffu;
      }
    }
  };
  __device__ Words cat_words(int set, int cat) const {
    const uint64_t *w = &L.nzm[set][cat_word(cat)];
    return Words{w[0], w[1], w[2], w[3]};
  }

  // the level at scan position i of block b of a category
  __device__ int blk_level(int cat, int b, int i) const {
    const ia_mbcode &c = L.code;
    int bx, by;
    block_pos(b, &bx, &by);
    switch (cat) {
      case 0: return c.dc[i];
      case 1: return c.luma[by * 4 + bx][i + 1];
      case 2: return c.luma[by * 4 + bx][i];
      case 3: return c.cdc[b][i];
      case 4: return c.cac[b >> 2][b & 3][i + 1];
      default: return c.luma8[b][i];
    }
  }

  // The slice parameters the rate count reads, loaded once per count (read through P inside its
  // branches, they were loaded from memory again in every branch)
  struct CabEnv {
    int p_slice, b_slice, transform8, nref;
  };
  // A category's first significance, last and level contexts (frame coding) as selections among
  // constants, so that a lane's category reads no table from memory
#define IA_CAT_SELECT(T)                                                                                \
  constexpr int v0 = T[0], v1 = T[1], v2 = T[2], v3 = T[3], v4 = T[4], v5 = T[5];                      \
  return cat == 0 ? v0 : cat == 1 ? v1 : cat == 2 ? v2 : cat == 3 ? v3 : cat == 4 ? v4 : v5;
  __device__ static int sig_off(int cat) { IA_CAT_SELECT(IA_SIG_OFF[0]) }
  __device__ static int last_off(int cat) { IA_CAT_SELECT(IA_LAST_OFF[0]) }
  __device__ static int abs_off(int cat) { IA_CAT_SELECT(IA_ABS_OFF) }
#undef IA_CAT_SELECT

  // The mb_type of a B macroblock (Table 7-14; enc/frame.rs:2343-2351)
  __device__ static int b_mb_type(const ia_mbcode &code) {
    const uint8_t PAIRS[9][2] = {{1, 1}, {2, 2}, {1, 2}, {2, 1}, {1, 3}, {2, 3}, {3, 1}, {3, 2}, {3, 3}};
    if (code.part == 0) return code.bdir[0];
    if (code.part == 3) return 22;
    int i = 0;
    while (i < 8 && !(PAIRS[i][0] == code.bdir[0] && PAIRS[i][1] == code.bdir[1])) i++;
    return 4 + 2 * i + (code.part == 2);
  }

  // mb_skip_flag, mb_type, transform_size_8x8_flag, intra_chroma_pred_mode, sub_mb_type, ref_idx,
  // coded_block_pattern and mb_qp_delta (enc/frame.rs:2028-2110, :2150-2200)
  __device__ uint32_t header_bins(uint8_t *st, const CabEnv &E) const {
    const ia_mbcode &code = L.code;
    const MbState *left = nb(-1, 0), *top = nb(0, -1);
    const int kind = code.kind;
    uint32_t bits = 0;
    if (E.p_slice) {
      dec_at(st, (E.b_slice ? 24 : 11) + (left && !left->skip) + (top && !top->skip), code.skip, bits);
      if (code.skip) return bits;
    }
    const int cbpl = code.cbp & 15, cbpc = code.cbp >> 4, prefix_ok = E.p_slice;
    const int base = E.b_slice ? 32 : E.p_slice ? 17 : 3;
    const int b_ctx = 27 + (left && !left->bdirect) + (top && !top->bdirect);
    if (kind == K_INTER && E.b_slice) {
      // B_Direct_16x16 0, B_L0_16x16 100, B_L1_16x16 101, the others 11 and four or five more bins
      // (Table 9-37); sub_mb_type B_Direct_8x8 0, B_L0_8x8 100, B_L1_8x8 101, B_Bi_8x8 11000
      const int t = b_mb_type(code);
      dec_at(st, b_ctx, t != 0, bits);
      if (t == 1 || t == 2) {
        dec_at(st, 30, 0, bits);
        dec_at(st, 32, t == 2, bits);
      } else if (t != 0) {
        dec_at(st, 30, 1, bits);
        const int v = t >= 3 && t <= 10 ? t - 3 : t == 11 ? 14 : t == 22 ? 15 : (t + 4) >> 1;
        dec_at(st, 31, (v >> 3) & 1, bits);
        for (int sh = 2; sh >= 0; sh--) dec_at(st, 32, (v >> sh) & 1, bits);
        if (t >= 12 && t <= 21) dec_at(st, 32, (t + 4) & 1, bits);
      }
      if (t == 22)
        for (int k = 0; k < 4; k++) {
          const int d = code.bdir[k];
          dec_at(st, 36, d != 0, bits);
          if (d == 1 || d == 2) {
            dec_at(st, 37, 0, bits);
            dec_at(st, 39, d == 2, bits);
          } else if (d == 3) {
            dec_at(st, 37, 1, bits);
            dec_at(st, 38, 0, bits);
            dec_at(st, 39, 0, bits);
            dec_at(st, 39, 0, bits);
          }
        }
    } else if (kind == K_INTER) {
      const int b1 = code.part == 1 || code.part == 2, b2 = code.part == 1 || code.part == 3;
      dec_at(st, 14, 0, bits);
      dec_at(st, 15, b1, bits);
      dec_at(st, b1 ? 17 : 16, b2, bits);
    } else {
      int first;
      if (E.b_slice) {
        // 111101: an intra type follows, with the suffix contexts
        dec_at(st, b_ctx, 1, bits);
        dec_at(st, 30, 1, bits);
        dec_at(st, 31, 1, bits);
        dec_at(st, 32, 1, bits);
        dec_at(st, 32, 0, bits);
        dec_at(st, 32, 1, bits);
        first = 32;
      } else if (prefix_ok) {
        dec_at(st, 14, 1, bits);
        first = 17;
      } else {
        first = 3 + (left && !(left->kind == K_I4 || left->kind == K_I8)) + (top && !(top->kind == K_I4 || top->kind == K_I8));
      }
      if (kind == K_I4 || kind == K_I8) {
        dec_at(st, first, 0, bits);
        if (E.transform8) dec_at(st, 399 + (left && left->t8x8) + (top && top->t8x8), kind == K_I8, bits);
      } else {
        dec_at(st, first, 1, bits);
        bits += 1;  // terminate bin 0
        const int b1 = prefix_ok ? base + 1 : base + 3, b2 = prefix_ok ? base + 2 : base + 4, b3 = prefix_ok ? base + 3 : base + 6;
        dec_at(st, b1, cbpl != 0, bits);
        dec_at(st, b2, cbpc != 0, bits);
        if (cbpc != 0) dec_at(st, prefix_ok ? base + 2 : base + 5, cbpc == 2, bits);
        dec_at(st, b3, (code.i16_mode >> 1) == 1, bits);
        dec_at(st, prefix_ok ? base + 3 : base + 7, (code.i16_mode & 1) == 1, bits);
      }
    }
    if (kind != K_INTER) {
      const int m = code.chroma_mode;
      const int ca = left && !left->inter && left->chroma_mode != 0, cb = top && !top->inter && top->chroma_mode != 0;
      dec_at(st, 64 + ca + cb, m > 0, bits);
      if (m > 0) {
        dec_at(st, 67, m > 1, bits);
        if (m > 1) dec_at(st, 67, m > 2, bits);
      }
    } else {
      if (!E.b_slice && code.part == 3) {
        for (int k = 0; k < 4; k++) {
          const int s = code.sub[k];
          dec_at(st, 21, s == 0, bits);
          if (s != 0) {
            dec_at(st, 22, s != 1, bits);
            if (s != 1) dec_at(st, 23, s == 2, bits);
          }
        }
      }
      if (!E.b_slice && E.nref > 1) {
        uint32_t grid = 0;  // reference index per 8x8 block so far, a byte each
        const int np = code.part == 3 ? 4 : NPARTS[code.part];
        for (int k = 0; k < np; k++) {
          int bx, by, w, h;
          if (code.part == 3) { bx = (k & 1) * 2; by = (k >> 1) * 2; w = h = 2; }
          else { bx = PARTS[code.part][k][0]; by = PARTS[code.part][k][1]; w = PARTS[code.part][k][2]; h = PARTS[code.part][k][3]; }
          const int a = bx > 0 ? ((grid >> (8 * ((by / 2) * 2 + (bx - 1) / 2))) & 0xff) > 0 : (left && left->refs[(by / 2) * 2 + 1] > 0);
          const int bb = by > 0 ? ((grid >> (8 * (((by - 1) / 2) * 2 + bx / 2))) & 0xff) > 0 : (top && top->refs[2 + bx / 2] > 0);
          const int r = code.refs[k], cbase = 54 + a + 2 * bb;
          for (int i = 0; i <= r; i++) dec_at(st, i == 0 ? cbase : (i == 1 ? 58 : 59), i < r, bits);
          for (int y8 = by / 2; y8 <= (by + h - 1) / 2; y8++)
            for (int x8 = bx / 2; x8 <= (bx + w - 1) / 2; x8++) {
              const int sh = 8 * (y8 * 2 + x8);
              grid = (grid & ~(0xffu << sh)) | ((uint32_t)r << sh);
            }
        }
      }
    }
    if (kind != K_I16) {
      int luma = 0;
      for (int b8 = 0; b8 < 4; b8++) {
        const int cond_a = (b8 & 1) ? ((luma >> (b8 - 1)) & 1) == 0 : (left && ((left->cbp >> (b8 + 1)) & 1) == 0);
        const int cond_b = b8 >= 2 ? ((luma >> (b8 - 2)) & 1) == 0 : (top && ((top->cbp >> (b8 + 2)) & 1) == 0);
        const int bit = (cbpl >> b8) & 1;
        dec_at(st, 73 + cond_a + 2 * cond_b, bit, bits);
        luma |= bit << b8;
      }
      int ca = left && (left->cbp >> 4) != 0, cb = top && (top->cbp >> 4) != 0;
      dec_at(st, 77 + ca + 2 * cb, cbpc != 0, bits);
      if (cbpc != 0) {
        ca = left && (left->cbp >> 4) == 2;
        cb = top && (top->cbp >> 4) == 2;
        dec_at(st, 81 + ca + 2 * cb, cbpc == 2, bits);
      }
    }
    if (kind == K_INTER && E.transform8 && cbpl != 0 && (code.part != 3 || (code.sub[0] | code.sub[1] | code.sub[2] | code.sub[3]) == 0))
      dec_at(st, 399 + (left && left->t8x8) + (top && top->t8x8), code.t8x8, bits);
    if (code.cbp != 0 || kind == K_I16) dec_at(st, 60, 0, bits);
    return bits;
  }

  // prev_intra_pred_mode_flag and rem_intra_pred_mode of the Intra4x4/8x8 blocks (contexts 68, 69);
  // the modes coded so far four bits each in a register
  __device__ uint32_t mode_bins(uint8_t *st) const {
    const ia_mbcode &code = L.code;
    if (code.kind != K_I4 && code.kind != K_I8) return 0;
    const MbState *left = nb(-1, 0), *top = nb(0, -1);
    const int t8 = code.kind == K_I8;
    uint32_t s68 = st[68], s69 = st[69], bits = 0;
    uint64_t so_far = 0x2222222222222222ull;
    for (int n = 0; n < (t8 ? 4 : 16); n++) {
      int bx, by;
      if (t8) { bx = (n & 1) * 2; by = (n >> 1) * 2; }
      else block_pos(n, &bx, &by);
      const int mode = code.modes[by * 4 + bx];
      int a = -1, b = -1;
      if (bx > 0) a = (int)((so_far >> (4 * (by * 4 + bx - 1))) & 15);
      else if (left) a = left->i4 ? left->modes[by * 4 + 3] : 2;
      if (by > 0) b = (int)((so_far >> (4 * ((by - 1) * 4 + bx))) & 15);
      else if (top) b = top->i4 ? top->modes[12 + bx] : 2;
      const int pred = a >= 0 && b >= 0 ? min(a, b) : 2;
      const uint64_t m = (uint64_t)mode;
      const int i = by * 4 + bx;
      so_far = (so_far & ~(15ull << (4 * i))) | (m << (4 * i));
      if (t8) {
        so_far = (so_far & ~(15ull << (4 * (i + 1)))) | (m << (4 * (i + 1)));
        so_far = (so_far & ~(15ull << (4 * (i + 4)))) | (m << (4 * (i + 4)));
        so_far = (so_far & ~(15ull << (4 * (i + 5)))) | (m << (4 * (i + 5)));
      }
      dec(s68, mode == pred, bits);
      if (mode != pred) {
        const int rem = mode < pred ? mode : mode - 1;
        for (int k = 0; k < 3; k++) dec(s69, (rem >> k) & 1, bits);
      }
    }
    st[68] = (uint8_t)s68;
    st[69] = (uint8_t)s69;
    return bits;
  }

  // one component of the motion vector differences, partitions in coding order, list 0's then list 1's
  // in a B slice (enc/frame.rs:2110-2150; the two lists share the component's contexts); min(|mvd|, 64)
  // of the 4x4 blocks coded so far a byte each in two registers
  __device__ uint32_t mvd_comp_bins(uint8_t *st, int comp, const CabEnv &E) const {
    const ia_mbcode &code = L.code;
    if (code.kind != K_INTER || code.skip) return 0;
    const MbState *left = nb(-1, 0), *top = nb(0, -1);
    const int base = comp == 0 ? 40 : 47, part = code.part, bsl = E.b_slice;
    uint32_t bits = 0;
    for (int l = 0; l < (bsl ? 2 : 1); l++) {
    uint64_t g[2] = {0, 0};
    int k = 0;
    for (int q = 0; q < (part == 3 && !bsl ? 4 : NPARTS[part]); q++) {
      const int sub = part == 3 && !bsl ? code.sub[q] : 0;
      for (int s = 0; s < (part == 3 && !bsl ? NSUBS[sub] : 1); s++, k++) {
        if (bsl && !((code.bdir[q] >> l) & 1)) continue;
        int bx, by, w, h;
        if (part == 3 && !bsl) {
          bx = (q & 1) * 2 + SUBS[sub][s][0];
          by = (q >> 1) * 2 + SUBS[sub][s][1];
          w = SUBS[sub][s][2];
          h = SUBS[sub][s][3];
        } else {
          bx = PARTS[part][q][0];
          by = PARTS[part][q][1];
          w = PARTS[part][q][2];
          h = PARTS[part][q][3];
        }
        auto at = [&](int i) { return (int)((g[i >> 3] >> (8 * (i & 7))) & 0xff); };
        const uint8_t(*lm)[2] = left ? (l ? left->mvd1 : left->mvd) : nullptr;
        const uint8_t(*tm)[2] = top ? (l ? top->mvd1 : top->mvd) : nullptr;
        const int a = bx > 0 ? at(by * 4 + bx - 1) : (lm ? lm[by * 4 + 3][comp] : 0);
        const int bb = by > 0 ? at((by - 1) * 4 + bx) : (tm ? tm[12 + bx][comp] : 0);
        const uint32_t sum = (uint32_t)(a + bb);
        const int d = bsl ? code.bmvd[l][k][comp] : code.mvds[k][comp];
        const uint32_t ad = (uint32_t)iabs(d);
        dec_at(st, base + (sum < 3 ? 0 : (sum > 32 ? 2 : 1)), ad > 0, bits);
        if (ad > 0) {
          int ctx = base + 3;
          for (uint32_t t = 1; t < (ad < 9 ? ad : 9); t++) {
            dec_at(st, ctx, 1, bits);
            if (ctx < base + 6) ctx++;
          }
          if (ad < 9) dec_at(st, ctx, 0, bits);
          else bits += 256u * (uint32_t)eg_bins(ad - 9, 3);
          bits += 256;  // sign
        }
        const uint64_t v = (uint64_t)min((int)ad, 64);
        for (int y = by; y < by + h; y++)
          for (int x = bx; x < bx + w; x++) {
            const int i = y * 4 + x, sh = 8 * (i & 7);
            if (i < 8) g[0] = (g[0] & ~(0xffull << sh)) | (v << sh);
            else g[1] = (g[1] & ~(0xffull << sh)) | (v << sh);
          }
      }
    }
    }
    return bits;
  }

  // coded_block_flag of each coded block of a category (contexts 85 + 4 cat + a + 2 b), the four
  // states packed in a register
  __device__ uint32_t cbf_bins(uint8_t *st, int cat) const {
    const ia_mbcode &code = L.code;
    const MbState *left = nb(-1, 0), *top = nb(0, -1);
    const int un = code.kind != K_INTER, cbpl = code.cbp & 15, ctx0 = 85 + 4 * cat;
    const Words nzw = cat_words(0, cat);
    uint32_t pk = (uint32_t)st[ctx0] | ((uint32_t)st[ctx0 + 1] << 8) | ((uint32_t)st[ctx0 + 2] << 16) | ((uint32_t)st[ctx0 + 3] << 24);
    uint32_t bits = 0;
    for (int b = 0; b < cat_blocks(cat); b++) {
      int a, bb;
      if (cat == 0) {
        a = left ? (left->cbf_dc & 1) != 0 : un;
        bb = top ? (top->cbf_dc & 1) != 0 : un;
      } else if (cat == 1 || cat == 2) {
        if (code.t8x8 || !((cbpl >> (b >> 2)) & 1)) continue;
        int bx, by;
        block_pos(b, &bx, &by);
        a = bx > 0 ? code.nnz[by * 4 + bx - 1] > 0 : (left ? left->nnz[by * 4 + 3] > 0 : un);
        bb = by > 0 ? code.nnz[(by - 1) * 4 + bx] > 0 : (top ? top->nnz[12 + bx] > 0 : un);
      } else if (cat == 3) {
        a = left ? (left->cbf_dc & (2 << b)) != 0 : un;
        bb = top ? (top->cbf_dc & (2 << b)) != 0 : un;
      } else {
        const int comp = b >> 2, bx = b & 1, by = (b >> 1) & 1;
        const uint8_t *cur = code.cnnz[comp];
        a = bx > 0 ? cur[by * 2] > 0 : (left ? left->cnnz[comp][by * 2 + 1] > 0 : un);
        bb = by > 0 ? cur[bx] > 0 : (top ? top->cnnz[comp][2 + bx] > 0 : un);
      }
      const int sh = 8 * (a + 2 * bb);
      uint32_t s = (pk >> sh) & 0xffu;
      dec(s, nzw.block(cat, b) != 0, bits);
      pk = (pk & ~(0xffu << sh)) | (s << sh);
    }
    for (int i = 0; i < 4; i++) st[ctx0 + i] = (uint8_t)(pk >> (8 * i));
    return bits;
  }

  // significance map lane j of a category: positions j of every coded block (significant_coeff_flag
  // and last_significant_coeff_flag on contexts sig + j, last + j); for 8x8 blocks the positions of
  // significance context j (j < 15) or of last context j - 15
  __device__ uint32_t sig_bins(uint8_t *st, int cat, int j) const {
    const int max = cat_max(cat);
    uint64_t smask, tmask;
    int sctx, lctx;
    if (cat == 5) {
      smask = j < 15 ? L.pos8[j] : 0;
      tmask = j < 15 ? 0 : L.pos8[j];
      sctx = IA_SIG_OFF[0][5] + j;
      lctx = IA_LAST_OFF[0][5] + j - 15;
    } else {
      smask = tmask = 1ull << j;
      sctx = sig_off(cat) + j;
      lctx = last_off(cat) + j;
    }
    const Words nzw = cat_words(0, cat);
    uint32_t ss = smask ? st[sctx] : 0, ts = tmask ? st[lctx] : 0, bits = 0;
    for (int b = 0; b < cat_blocks(cat); b++) {
      const uint64_t m = nzw.block(cat, b);
      if (m == 0) continue;
      const int last = 63 - __builtin_clzll(m);
      uint64_t q = (smask | tmask) & ((2ull << min(last, max - 2)) - 1);
      while (q) {
        const int p = __builtin_ctzll(q);
        q &= q - 1;
        const int nzp = (int)((m >> p) & 1);
        if ((smask >> p) & 1) dec(ss, nzp, bits);
        if (((tmask >> p) & 1) && nzp) dec(ts, p == last, bits);
      }
    }
    if (smask) st[sctx] = (uint8_t)ss;
    if (tmask) st[lctx] = (uint8_t)ts;
    return bits;
  }

  // Level context k of a category (coeff_abs_level_minus1). The levels of a block are coded from the
  // last to the first; while no level above 1 has been coded, the r-th takes its first bin on context
  // min(1 + r, 4), after the first such level on context 0; the first bin is 1 for the levels above 1
  // only. The j-th level above 1 codes its further bins on context 5 + min(j, 4) (5 + min(j, 3) for
  // chroma DC). So the masks of nonzero levels and of levels above 1 give each context its bins; only
  // the levels above 1 are read. Lane k = 0 also counts the bypass bins of the signs, the lanes of the
  // further bins the Exp-Golomb suffixes of their levels.
  __device__ uint32_t level_bins(uint8_t *st, int cat, int k) const {
    const int ctx = abs_off(cat) + k, gmax = 4 - (cat == 3);
    const Words nzw = cat_words(0, cat), g1w = cat_words(1, cat);
    uint32_t s = st[ctx], bits = 0;
    for (int b = 0; b < cat_blocks(cat); b++) {
      const uint64_t m = nzw.block(cat, b);
      if (m == 0) continue;
      const uint64_t g = g1w.block(cat, b);
      const int n = __popcll(m);
      // levels coded before the first level above 1 (all of them without one)
      const int hg = g ? 63 - __builtin_clzll(g) : -1;
      const int r1 = g ? __popcll(m >> hg) - 1 : n;  // its index from the last level
      if (k == 0) {
        bits += 256u * (uint32_t)n;  // signs
        // the levels after the first level above 1, from the last to the first
        for (uint64_t q = g ? m & ((1ull << hg) - 1) : 0; q != 0;) {
          const int i = 63 - __builtin_clzll(q);
          q &= ~(1ull << i);
          dec(s, (int)((g >> i) & 1), bits);
        }
      } else if (k < 4) {
        if (k - 1 < n && k - 1 <= r1) dec(s, k - 1 == r1, bits);
      } else if (k == 4) {
        for (int r = 3; r < min(r1, n); r++) dec(s, 0, bits);
        if (r1 >= 3 && r1 < n) dec(s, 1, bits);
      } else {
        // the levels above 1 whose further bins use this context: the (k - 5)-th, or from the
        // gmax-th on
        int j = 0;
        for (uint64_t q = g; q != 0; j++) {
          const int i = 63 - __builtin_clzll(q);
          q &= ~(1ull << i);
          if (5 + min(j, gmax) != k) continue;
          const int a = iabs(blk_level(cat, b, i));
          for (int t = 0; t < min(a - 2, 13); t++) dec(s, 1, bits);
          if (a < 15) dec(s, 0, bits);
          else bits += 256u * (uint32_t)eg_bins((uint32_t)(a - 15), 0);
        }
      }
    }
    st[ctx] = (uint8_t)s;
    return bits;
  }

  // The rate of the macroblock in L.code in 1/256 bits, coded on the context states st (which it
  // leaves as the writer would); the same in every lane. The chroma residual's bins (coded_block_flags
  // and levels of the chroma DC and AC blocks) are counted when *chroma < 0, their cost then stored
  // there; otherwise *chroma is taken as their cost (the same chroma coded before on the same
  // states) and their contexts are left alone.
  __device__ uint32_t count_bins(uint8_t *st, int *chroma) {
    nz_masks();
    sync();
    return count_masked(st, chroma);
  }
  // the same, the masks of nz_masks already in L.nzm
  __device__ uint32_t count_masked(uint8_t *st, int *chroma) {
    const ia_mbcode &code = L.code;
    const int cbpl = code.cbp & 15, cbpc = code.cbp >> 4, i16 = code.kind == K_I16, t8 = code.t8x8;
    const int count_chroma = *chroma < 0;
    // the categories coded, in a fixed order, four bits each
    uint32_t cats = 0;
    int ncat = 0, total = 4;
    auto add = [&](int cat) {
      cats |= (uint32_t)cat << (4 * ncat++);
      total += cat_jobs(cat);
    };
    if (i16) add(0);
    if (cbpl && !t8) add(i16 ? 1 : 2);
    if (cbpl && t8) add(5);
    if (count_chroma && cbpc > 0) add(3);
    if (count_chroma && cbpc == 2) add(4);
    const CabEnv E{__builtin_amdgcn_readfirstlane(P.p_slice), __builtin_amdgcn_readfirstlane(P.b_slice),
                   __builtin_amdgcn_readfirstlane(P.transform8), __builtin_amdgcn_readfirstlane(P.nref)};
    uint32_t bits = 0, cbits = 0;
    for (int j = lane(); j < total; j += 64) {
      if (j == 0) bits += header_bins(st, E);
      else if (j == 1) bits += mode_bins(st);
      else if (j < 4) bits += mvd_comp_bins(st, j - 2, E);
      else {
        int t = j - 4, ci = 0;
        while (t >= cat_jobs((cats >> (4 * ci)) & 15)) t -= cat_jobs((cats >> (4 * ci++)) & 15);
        const int cat = (cats >> (4 * ci)) & 15;
        uint32_t b;
        if (cat != 5 && t == 0) {
          b = cbf_bins(st, cat);
        } else {
          t -= cat != 5;
          if (t < cat_sig_jobs(cat)) b = sig_bins(st, cat, t);
          else b = level_bins(st, cat, t - cat_sig_jobs(cat));
        }
        bits += b;
        if (cat == 3 || cat == 4) cbits += b;
      }
    }
    uint32_t total_bits = (uint32_t)wave_sum((int)bits);
    if (count_chroma) *chroma = wave_sum((int)cbits);
    else total_bits += (uint32_t)*chroma;
    sync();
    return total_bits;
  }

  __device__ static void mvd_grid(const Rect *rs, int n, const int16_t (*mvds)[2], uint8_t (*out)[2]) {
    for (int i = 0; i < 16; i++) out[i][0] = out[i][1] = 0;
    for (int k = 0; k < n; k++)
      for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
        for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++)
          for (int q = 0; q < 2; q++) out[by * 4 + bx][q] = (uint8_t)min(iabs(mvds[k][q]), 64);
  }

  // J of the candidate in L.code whose reconstruction is in the canvas (enc/frame.rs:1453-1457)
  // *chroma: the chroma residual's rate when known (see count_bins), else -1
  // The chroma residual's contexts, 61 in four runs: the coded_block_flags of chroma DC and AC
  // (85 + 4 cat, cats 3 and 4), then for each of significance, last and level the chroma DC
  // contexts followed by the chroma AC ones.
  __device__ static int chroma_ctx(int i) {
    return i < 8 ? 85 + 12 + i : i < 25 ? IA_SIG_OFF[0][3] + i - 8 : i < 42 ? IA_LAST_OFF[0][3] + i - 25 : IA_ABS_OFF[3] + i - 42;
  }

  // J of the candidate in L.code; its rate in *bits and the context states after its bins in L.tmp.
  // When its chroma is that of the last full count (*chroma >= 0), the chroma contexts start at the
  // states that count left them in, so L.tmp still ends as the writer's would.
  // (The masks of nz_masks are in L.nzm and ssd is the candidate's distortion.)
  __device__ double rd_cost(double lam2, int *chroma, uint32_t *bits, long long ssd) {
    const int full = *chroma < 0;
    {
      const uint32_t *s = reinterpret_cast<const uint32_t *>(L.cabac);
      uint32_t *d = reinterpret_cast<uint32_t *>(L.tmp);
      for (int i = lane(); i < 256; i += 64) d[i] = s[i];
    }
    sync();
    if (!full && lane() < 61) L.tmp[chroma_ctx(lane())] = L.chroma_st[lane()];
    sync();
    *bits = count_masked(L.tmp, chroma);
    if (full && lane() < 61) L.chroma_st[lane()] = L.tmp[chroma_ctx(lane())];
    sync();
    return (double)ssd + lam2 * ((double)*bits / 256.0);
  }

  // A floor of the rate (1/256 bits) of the candidate whose masks nz_masks left in L.nzm: one bypass
  // bin per nonzero level (its sign), the chroma levels' replaced by the chroma's rate when it is
  // known (chroma >= 0). The same in every lane.
  __device__ uint32_t rate_floor(int chroma) const {
    uint32_t n = 0;
    for (int w = 0; w < 8; w++) n += (uint32_t)__popcll(L.nzm[0][w]);  // luma 4x4 and 8x8 blocks
    const uint64_t w10 = L.nzm[0][10];
    n += (uint32_t)__popcll(w10 & 0xffffull);  // Intra16x16 DC
    if (chroma >= 0) return 256u * n + (uint32_t)chroma;
    n += (uint32_t)(__popcll(L.nzm[0][8]) + __popcll(L.nzm[0][9]) + __popcll(w10 >> 16));
    return 256u * n;
  }

  __device__ void copy_code(ia_mbcode *dst, const ia_mbcode *src) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst);
    for (int i = lane(); i < (int)(sizeof(ia_mbcode) / 4); i += 64) d[i] = s[i];
  }

  // the candidate just costed replaces the kept one when it is the first or its J is lower: its code,
  // reconstruction, rate and final context states
  __device__ void keep(double j, uint32_t bits, int *have, double *best_j) {
    if (!*have || j < *best_j) {
      *have = 1;
      *best_j = j;
      copy_code(&L.keep, &L.code);
      save_rec(L.keep_rec);
      const uint32_t *s = reinterpret_cast<const uint32_t *>(L.tmp);
      uint32_t *d = reinterpret_cast<uint32_t *>(L.kept_st);
      for (int i = lane(); i < 110; i += 64) d[i] = s[i];
      if (lane() == 0) L.kept_bits = bits;
    }
    sync();
  }

  // The chroma of a candidate (mode, levels, nonzero counts, coded pattern) kept in L.csave: every
  // intra candidate codes the same chroma (enc/frame.rs:1807-1852 depends on the chroma samples and
  // edges only), so the candidates after the first take it from there; their reconstructed chroma is
  // already in the canvas.
  static_assert(offsetof(ia_mbcode, cac) == offsetof(ia_mbcode, cdc) + 16 && offsetof(ia_mbcode, cnnz) % 4 == 0 &&
                offsetof(ia_mbcode, cdc) % 4 == 0, "the chroma levels as words, DC first");
  __device__ void save_chroma() {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&L.code);
    uint32_t *d = reinterpret_cast<uint32_t *>(&L.csave);
    constexpr int cnnz = offsetof(ia_mbcode, cnnz) / 4, cdc = offsetof(ia_mbcode, cdc) / 4;
    // cnnz: 2 words; cdc and cac: 4 + 64 words
    if (lane() < 2) d[lane()] = s[cnnz + lane()];
    for (int i = lane(); i < 68; i += 64) d[2 + i] = s[cdc + i];
    if (lane() == 0) {
      L.csave.mode = L.code.chroma_mode;
      L.csave.cbp = L.code.cbp & 0xf0;
    }
    sync();
  }
  __device__ void restore_chroma() {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&L.csave);
    uint32_t *d = reinterpret_cast<uint32_t *>(&L.code);
    constexpr int cnnz = offsetof(ia_mbcode, cnnz) / 4, cdc = offsetof(ia_mbcode, cdc) / 4;
    if (lane() < 2) d[cnnz + lane()] = s[lane()];
    for (int i = lane(); i < 68; i += 64) d[cdc + i] = s[2 + i];
    if (lane() == 0) {
      L.code.chroma_mode = L.csave.mode;
      L.code.cbp = (uint8_t)((L.code.cbp & 15) | L.csave.cbp);
    }
    sync();
  }

  // The inter candidate in L.code again with the 8x8 transform: its vectors, prediction (L.pred) and
  // chroma stay those of the 4x4 transform's (enc/frame.rs:1201-1252 predicts and codes the chroma
  // alike for both), so only the luma residual is coded again.
  __device__ void recode_inter_t8() {
    uint32_t *lv = reinterpret_cast<uint32_t *>(&L.code.luma[0][0]);
    uint32_t *lv8 = reinterpret_cast<uint32_t *>(&L.code.luma8[0][0]);
    for (int i = lane(); i < 128; i += 64) {
      lv[i] = 0;
      lv8[i] = 0;
    }
    if (lane() < 16) L.code.nnz[lane()] = 0;
    if (lane() == 0) {
      L.code.cbp &= 0xf0;
      L.code.t8x8 = 0;
    }
    sync();
    code_luma8();
  }

  // Costs the candidate in L.code and keeps it when its J is the least so far. Its J is at least
  // its distortion plus lam2 times a floor of its rate (rate_floor; the same monotone double
  // arithmetic as J's); a candidate whose floor already reaches the kept J could not be kept, so its
  // bins are not counted (it leaves *chroma and the chroma states alone: the next candidate with
  // the same chroma counts them).
  __device__ void rd_keep(double lam2, int *chroma, int *have, double *best_j) {
    nz_masks();
    sync();
    const long long ssd = ssd_mb();
    if (*have && !((double)ssd + lam2 * ((double)rate_floor(*chroma) / 256.0) < *best_j)) {
      sync();
      return;
    }
    uint32_t b;
    const double j = rd_cost(lam2, chroma, &b, ssd);
    keep(j, b, have, best_j);
  }

  __device__ void decide_rd(int i16, const Inter *inter, ia_mbstats *st) {
    const double lam2 = P.lam2;
    int have = 0;
    double best_j = 0.0;
    if (inter) {
      int16_t sm[2];
      skip_mv(sm);
      int16_t mvs[16][2];
      for (int i = 0; i < 16; i++) { mvs[i][0] = sm[0]; mvs[i][1] = sm[1]; }
      const uint8_t z4[4] = {0, 0, 0, 0};
      code_new(K_INTER);
      if (lane() == 0) {
        L.code.skip = 1;
        for (int i = 0; i < 16; i++) { L.code.mvs[i][0] = sm[0]; L.code.mvs[i][1] = sm[1]; }
      }
      sync();
      inter_pred(0, z4, mvs, z4);
      for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
      for (int i = lane(); i < 128; i += 64) ccv(i >> 6, (i & 63) & 7, (i & 63) >> 3) = L.cpred[i >> 6][i & 63];
      sync();
      int none = -1;
      rd_keep(lam2, &none, &have, &best_j);
      int inter_chroma = -1;
      code_inter(inter->part, inter->mvs, inter->refs, inter->sub, 0);
      if (!(L.code.cbp == 0 && inter->part == 0 && inter->mvs[0][0] == sm[0] && inter->mvs[0][1] == sm[1] && inter->refs[0] == 0))
        rd_keep(lam2, &inter_chroma, &have, &best_j);
      if (P.transform8 && (inter->sub[0] | inter->sub[1] | inter->sub[2] | inter->sub[3]) == 0) {
        recode_inter_t8();
        if (L.code.t8x8) rd_keep(lam2, &inter_chroma, &have, &best_j);
      }
    }
    int intra_chroma = -1;
    code_i16(i16);
    code_intra_chroma();
    save_chroma();
    rd_keep(lam2, &intra_chroma, &have, &best_j);
    code_i4();
    restore_chroma();
    rd_keep(lam2, &intra_chroma, &have, &best_j);
    if (P.transform8) {
      code_i8();
      restore_chroma();
      rd_keep(lam2, &intra_chroma, &have, &best_j);
    }
    restore_rec(L.keep_rec);
    copy_code(&L.code, &L.keep);
    sync();
    st->rd_cost = best_j;
    rd_kept = 1;
  }

  // ---- B macroblocks (enc/frame.rs:1518-1780) ----
  //
  // A B slice predicts from one picture per list: P.hp/rcb/rcr [0] the list 0 picture, [1] the list 1
  // picture, P.rmbs0 the list 1 picture's records. Every partition a B macroblock considers moves as a
  // whole per 8x8 block, so a candidate's motion is held per 8x8 block (BMot) and spread over the 4x4
  // blocks only in the code.

  struct BMot {
    uint8_t use[4];         // per 8x8 block: bit 0 list 0, bit 1 list 1
    int16_t mv[2][4][2];    // per list and 8x8 block
  };
  struct BDir {
    int16_t mv[2][2];       // each list's direct vector
    uint8_t zero;           // bit 4 l + i8: list l's vector is zero in 8x8 block i8
    uint8_t use;            // lists used
  };

  __device__ static int min_positive(int x, int y) { return x >= 0 && y >= 0 ? min(x, y) : max(x, y); }

  // Spatial direct motion (enc/frame.rs:1717-1745, 8.4.1.2.2): per list the neighbours' smallest
  // reference index and the 16x16 predictor, zero in the 8x8 blocks whose co-located corner block of
  // the list 1 picture uses its reference 0 and moves by at most a quarter sample; both lists with zero
  // vectors when no neighbour uses either.
  __device__ void direct_mvs(BDir &d) const {
    const int16_t z[16][2] = {};
    const int8_t nr[16] = {};
    int refs[2];
    for (int l = 0; l < 2; l++) {
      int ra = -1, rb = -1, rc = -1, r;
      int16_t m[2];
      if (nb_motion(z, nr, 0, -1, 0, &r, m, l)) ra = r;
      if (nb_motion(z, nr, 0, 0, -1, &r, m, l)) rb = r;
      if (nb_motion(z, nr, 0, 4, -1, &r, m, l)) rc = r;
      else if (nb_motion(z, nr, 0, -1, -1, &r, m, l)) rc = r;
      refs[l] = min_positive(ra, min_positive(rb, rc));
    }
    d.mv[0][0] = d.mv[0][1] = d.mv[1][0] = d.mv[1][1] = 0;
    d.zero = 0;
    if (refs[0] < 0 && refs[1] < 0) {
      d.use = 3;
      return;
    }
    for (int l = 0; l < 2; l++)
      if (refs[l] >= 0) mv_pred(0, d.mv[l], l);
    const ia_mbinfo *col = P.rmbs0 ? &P.rmbs0[mby * P.mb_w + mbx] : nullptr;
    if (col && col->kind == K_INTER)
      for (int i8 = 0; i8 < 4; i8++) {
        const int cx = (i8 & 1) * 3, cy = (i8 >> 1) * 3;
        const int16_t *m = col->mv[cy * 4 + cx];
        if (col->ref_idx[i8] == 0 && iabs(m[0]) <= 1 && iabs(m[1]) <= 1)
          for (int l = 0; l < 2; l++)
            if (refs[l] == 0) d.zero |= (uint8_t)(1 << (4 * l + i8));
      }
    d.use = (uint8_t)((refs[0] >= 0) | ((refs[1] >= 0) << 1));
  }

  __device__ static void direct_mot(const BDir &d, BMot &m) {
    for (int i8 = 0; i8 < 4; i8++) {
      m.use[i8] = d.use;
      for (int l = 0; l < 2; l++) {
        const bool on = ((d.use >> l) & 1) && !((d.zero >> (4 * l + i8)) & 1);
        m.mv[l][i8][0] = on ? d.mv[l][0] : 0;
        m.mv[l][i8][1] = on ? d.mv[l][1] : 0;
      }
    }
  }

  // Eighth-sample chroma of the w x h block at chroma (cx, cy) from picture r into out (both
  // components, row stride 8, at (ox, oy)), the window clamped to the plane (h264/inter.rs:377-422).
  __device__ void chroma_pred(int r, int cx, int cy, int w, int h, int mx, int my, uint8_t (*out)[64], int ox, int oy) const {
    const int cw = P.mb_w * 8, ch = P.mb_h * 8;
    const int fx = mx & 7, fy = my & 7;
    const int x0 = cx + (mx >> 3), y0 = cy + (my >> 3);
    const int wa = (8 - fx) * (8 - fy), wb = fx * (8 - fy), wc = (8 - fx) * fy, wd = fx * fy;
    for (int i = lane(); i < 2 * w * h; i += 64) {
      const int ci = i / (w * h), j = i % (w * h), rr = j / w, c = j % w;
      const uint8_t *pl = ci == 0 ? P.rcb[r] : P.rcr[r];
      const int X = x0 + c, Y = y0 + rr;
#define S(xx, yy) ((int)pl[(size_t)clampi((yy), 0, ch - 1) * cw + clampi((xx), 0, cw - 1)])
      out[ci][(oy + rr) * 8 + ox + c] = (uint8_t)((wa * S(X, Y) + wb * S(X + 1, Y) + wc * S(X, Y + 1) + wd * S(X + 1, Y + 1) + 32) >> 6);
#undef S
    }
  }

  // The prediction of a B macroblock into L.pred / L.cpred, 8x8 block by 8x8 block from the lists used
  // there, averaged when both are (enc/frame.rs:1749-1780).
  __device__ void b_pred(const BMot &m) {
    for (int i8 = 0; i8 < 4; i8++) {
      const int bx = (i8 & 1) * 8, by = (i8 >> 1) * 8;
      for (int l = 0; l < 2; l++) {
        if (!((m.use[i8] >> l) & 1)) continue;
        uint8_t *dst = l ? L.bs.p1 : L.pred;
        luma_pred(l, mbx * 16 + bx, mby * 16 + by, 8, 8, m.mv[l][i8][0], m.mv[l][i8][1], dst + by * 16 + bx, 16);
        chroma_pred(l, mbx * 8 + bx / 2, mby * 8 + by / 2, 4, 4, m.mv[l][i8][0], m.mv[l][i8][1], l ? L.bs.c1 : L.cpred, bx / 2, by / 2);
      }
    }
    sync();
    for (int i = lane(); i < 384; i += 64) {
      const int c = i >= 256, j = c ? (i - 256) & 63 : i;
      const int i8 = c ? ((j >> 3) >> 2) * 2 + ((j & 7) >> 2) : ((j >> 4) >> 3) * 2 + ((j & 15) >> 3);
      const int u = m.use[i8];
      uint8_t *a = c ? &L.cpred[(i - 256) >> 6][j] : &L.pred[j];
      const int b = c ? L.bs.c1[(i - 256) >> 6][j] : L.bs.p1[j];
      if (u == 3) *a = (uint8_t)((*a + b + 1) >> 1);
      else if (u == 2) *a = (uint8_t)b;
    }
    sync();
  }

  // SATD of each 4x4 block (raster) of the macroblock's luma against pred (stride 16) into L.bs.sat
  __device__ void satd_blocks(const uint8_t *pred) {
    if (lane() < 16) {
      const int ox = (lane() & 3) * 4, oy = (lane() >> 2) * 4;
      int d[16];
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++) d[rr * 4 + c] = (int)L.src_y[(oy + rr) * 16 + ox + c] - (int)pred[(oy + rr) * 16 + ox + c];
      L.bs.sat[lane()] = satd4(d);
    }
    sync();
  }
  // the sum of L.bs.sat over the 4x4 blocks of the rectangle (bx, by, w, h) (4x4 units)
  __device__ static int sat_sum(int bx, int by, int w, int h) {
    int t = 0;
    for (int y = by; y < by + h; y++)
      for (int x = bx; x < bx + w; x++) t += L.bs.sat[y * 4 + x];
    return t;
  }

  // The two lists' predictions of partitions (bx[k], by[k], w, h), k < n, averaged into L.pred, and the
  // SATD of each 4x4 block against it in L.bs.sat (the bi-predicted candidates of
  // enc/frame.rs:1578-1581, :1637-1639)
  __device__ void bi_satd(int n, const int *bx, const int *by, int w, int h, const int16_t (*mv0)[2], const int16_t (*mv1)[2]) {
    for (int k = 0; k < n; k++) {
      luma_pred(0, mbx * 16 + bx[k] * 4, mby * 16 + by[k] * 4, w * 4, h * 4, mv0[k][0], mv0[k][1], L.pred + by[k] * 64 + bx[k] * 4, 16);
      luma_pred(1, mbx * 16 + bx[k] * 4, mby * 16 + by[k] * 4, w * 4, h * 4, mv1[k][0], mv1[k][1], L.bs.p1 + by[k] * 64 + bx[k] * 4, 16);
    }
    sync();
    for (int i = lane(); i < 256; i += 64) L.pred[i] = (uint8_t)((L.pred[i] + L.bs.p1[i] + 1) >> 1);
    sync();
    satd_blocks(L.pred);
  }

  // Each group's search of one partition in one list; group g searches (bx[g], by[g]) in list lst[g]
  // with the predictor of its list and the extra candidates of its list, the results read by every
  // lane (the groups' rectangles have one size).
  __device__ void b_search4(const int *bx, const int *by, int w, int h, const int *lst, const int16_t (*mvps)[2],
                            const int16_t (*extra)[5][2], int nextra, int16_t (*mv)[2], int *cost) {
    const int g = group();
    int16_t out[2];
    const int c = motion_search(0, bx[g], by[g], w, h, mvps[lst[g]], extra[lst[g]], nextra, out, lst[g]);
    if (glane() == 0) {
      L.bs.mv[g][0] = out[0];
      L.bs.mv[g][1] = out[1];
      L.bs.cost[g] = c;
    }
    sync();
    for (int k = 0; k < 4; k++) {
      mv[k][0] = L.bs.mv[k][0];
      mv[k][1] = L.bs.mv[k][1];
      cost[k] = L.bs.cost[k];
    }
    sync();
  }

  struct BParts {
    int part;
    uint8_t dirs[4];
    int16_t pmv[2][4][2];
  };

  // The best partitioning below 16x16 when it costs less than cost16 (enc/frame.rs:1621-1685): each
  // 8x8 block direct (d8: the direct prediction's SATD per 8x8 block) or from list 0, list 1 or both,
  // then 16x8 and 8x16 without direct. Costs are never negative, so a running total that reaches the
  // bound does so at the end too: every search runs (four at a time, one per group) and the totals are
  // compared once. Returns 0 when the 8x8 blocks cost cost16 or more.
  __device__ int best_b_parts(const int16_t (*mvs16)[2], const int16_t (*mvps)[2], int cost16, const int *d8, BParts &out) {
    const int lam = P.lambda;
    const int BX[4] = {0, 2, 0, 2}, BY[4] = {0, 0, 2, 2};
    int16_t extra[2][5][2];
    for (int l = 0; l < 2; l++) { extra[l][0][0] = mvs16[l][0]; extra[l][0][1] = mvs16[l][1]; }
    int16_t pm[2][4][2];
    int pc[2][4];
    for (int l = 0; l < 2; l++) {
      const int lst[4] = {l, l, l, l};
      b_search4(BX, BY, 2, 2, lst, mvps, extra, 1, pm[l], pc[l]);
    }
    bi_satd(4, BX, BY, 2, 2, pm[0], pm[1]);
    int total = lam * 6;
    for (int k = 0; k < 4; k++) {
      const int bi = sat_sum(BX[k], BY[k], 2, 2) + lam * (mvd_bits(pm[0][k][0], pm[0][k][1], mvps[0][0], mvps[0][1]) +
                                                           mvd_bits(pm[1][k][0], pm[1][k][1], mvps[1][0], mvps[1][1]) + 2);
      int bd, bc;
      if (bi < min(pc[0][k], pc[1][k])) { bd = 3; bc = bi; }
      else if (pc[0][k] <= pc[1][k]) { bd = 1; bc = pc[0][k]; }
      else { bd = 2; bc = pc[1][k]; }
      // sub_mb_type takes about one bin for direct and three for the others
      bc += lam * 3;
      if (d8[k] + lam < bc) { bd = 0; bc = d8[k] + lam; }
      out.dirs[k] = (uint8_t)bd;
      total += bc;
    }
    sync();
    if (total >= cost16) return 0;
    out.part = 3;
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 4; k++) { out.pmv[l][k][0] = pm[l][k][0]; out.pmv[l][k][1] = pm[l][k][1]; }
    int best_total = total;
    for (int l = 0; l < 2; l++) {
      for (int k = 0; k < 4; k++) { extra[l][k][0] = pm[l][k][0]; extra[l][k][1] = pm[l][k][1]; }
      extra[l][4][0] = mvs16[l][0];
      extra[l][4][1] = mvs16[l][1];
    }
    for (int part = 1; part < 3; part++) {
      // groups 0 and 1 the first partition in lists 0 and 1, groups 2 and 3 the second
      const int w = part == 1 ? 4 : 2, h = part == 1 ? 2 : 4;
      const int gx[4] = {0, 0, part == 2 ? 2 : 0, part == 2 ? 2 : 0}, gy[4] = {0, 0, part == 1 ? 2 : 0, part == 1 ? 2 : 0};
      const int lst[4] = {0, 1, 0, 1};
      int16_t m[4][2];
      int c[4];
      b_search4(gx, gy, w, h, lst, mvps, extra, 5, m, c);
      const int px[2] = {gx[0], gx[2]}, py[2] = {gy[0], gy[2]};
      const int16_t m0[2][2] = {{m[0][0], m[0][1]}, {m[2][0], m[2][1]}}, m1[2][2] = {{m[1][0], m[1][1]}, {m[3][0], m[3][1]}};
      bi_satd(2, px, py, w, h, m0, m1);
      int t2 = lam * 7;
      uint8_t d2[4] = {0, 0, 0, 0};
      for (int k = 0; k < 2; k++) {
        const int c0 = c[2 * k], c1 = c[2 * k + 1];
        const int bi = sat_sum(px[k], py[k], w, h) + lam * (mvd_bits(m0[k][0], m0[k][1], mvps[0][0], mvps[0][1]) +
                                                            mvd_bits(m1[k][0], m1[k][1], mvps[1][0], mvps[1][1]) + 2);
        int bd, bc;
        if (bi < min(c0, c1)) { bd = 3; bc = bi; }
        else if (c0 <= c1) { bd = 1; bc = c0; }
        else { bd = 2; bc = c1; }
        d2[k] = (uint8_t)bd;
        t2 += bc;
      }
      sync();
      if (t2 < best_total) {
        best_total = t2;
        out.part = part;
        for (int k = 0; k < 4; k++) out.dirs[k] = d2[k];
        for (int k = 0; k < 4; k++)
          for (int q = 0; q < 2; q++) {
            out.pmv[0][k][q] = k < 2 ? m0[k][q] : 0;
            out.pmv[1][k][q] = k < 2 ? m1[k][q] : 0;
          }
      }
    }
    return 1;
  }

  // The B fields of L.code for partitioning `part` with directions `dirs` and motion m (the vector
  // differences by enc/frame.rs:1688-1711: partitions predicted in order from the earlier ones using
  // the list, a direct 8x8 block counting with its derived motion), then its prediction and residual.
  __device__ void code_b(int part, const uint8_t *dirs, const BMot &m, int skip) {
    code_new(K_INTER);
    int16_t mvd[2][4][2] = {};
    if (!skip) {
      const int dir = part == 1 ? 1 : part == 2 ? 2 : 0;
      const int8_t nr[16] = {};
      for (int l = 0; l < 2; l++) {
        int16_t cur[16][2] = {};
        uint32_t done = 0;
        for (int k = 0; k < NPARTS[part]; k++) {
          const int x = PARTS[part][k][0], y = PARTS[part][k][1], w = PARTS[part][k][2], h = PARTS[part][k][3];
          const int i8 = (y / 2) * 2 + x / 2;
          if ((dirs[k] >> l) & 1) {
            int16_t p[2];
            pred_mv(cur, nr, done, x, y, w, dir, 0, p, l);
            mvd[l][k][0] = (int16_t)(m.mv[l][i8][0] - p[0]);
            mvd[l][k][1] = (int16_t)(m.mv[l][i8][1] - p[1]);
          }
          for (int yy = y; yy < y + h; yy++)
            for (int xx = x; xx < x + w; xx++) {
              const int b = yy * 4 + xx, b8 = (yy / 2) * 2 + xx / 2;
              cur[b][0] = m.mv[l][b8][0];
              cur[b][1] = m.mv[l][b8][1];
              done |= 1u << b;
              if (!((m.use[b8] >> l) & 1)) done |= 1u << (16 + b);
            }
        }
      }
    }
    if (lane() == 0) {
      L.code.skip = (uint8_t)skip;
      L.code.part = (uint8_t)part;
      for (int k = 0; k < 4; k++) {
        L.code.bdir[k] = dirs[k];
        L.code.buse[k] = m.use[k];
      }
      for (int l = 0; l < 2; l++) {
        for (int b = 0; b < 16; b++) {
          const int b8 = (b / 8) * 2 + (b % 4) / 2, on = (m.use[b8] >> l) & 1;
          L.code.bmv[l][b][0] = on ? m.mv[l][b8][0] : 0;
          L.code.bmv[l][b][1] = on ? m.mv[l][b8][1] : 0;
        }
        for (int k = 0; k < 4; k++) {
          L.code.bmvd[l][k][0] = mvd[l][k][0];
          L.code.bmvd[l][k][1] = mvd[l][k][1];
        }
      }
    }
    sync();
  }

  // A B macroblock by rate-distortion cost among B_Skip and B_Direct_16x16, the best of list 0, list 1
  // and both at 16x16 and the best smaller partitioning, Intra16x16 and Intra4x4 (enc/frame.rs:1518-1617)
  __device__ void decide_b(ia_mbstats *st) {
    const int lam = P.lambda;
    const double lam2 = P.lam2;
    int have = 0;
    double best_j = 0.0;
    BDir dd;
    direct_mvs(dd);
    BMot dm;
    direct_mot(dd, dm);
    b_pred(dm);
    // the direct prediction's SATD per 8x8 block (for B_Direct_8x8 in best_b_parts)
    satd_blocks(L.pred);
    int d8[4];
    for (int k = 0; k < 4; k++) d8[k] = sat_sum((k & 1) * 2, (k >> 1) * 2, 2, 2);
    sync();
    const uint8_t zdir[4] = {0, 0, 0, 0};
    // B_Skip: the direct prediction, nothing coded
    code_b(0, zdir, dm, 1);
    for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
    for (int i = lane(); i < 128; i += 64) ccv(i >> 6, (i & 63) & 7, (i & 63) >> 3) = L.cpred[i >> 6][i & 63];
    sync();
    int none = -1;
    rd_keep(lam2, &none, &have, &best_j);
    // B_Direct_16x16 with a residual (without one it is B_Skip at a higher price)
    int dchroma = -1;
    code_b(0, zdir, dm, 0);
    code_inter_residual(0);
    if (L.code.cbp != 0) rd_keep(lam2, &dchroma, &have, &best_j);
    if (P.transform8) {
      recode_inter_t8();
      if (L.code.cbp != 0 && L.code.t8x8) rd_keep(lam2, &dchroma, &have, &best_j);
    }
    // 16x16 from either list or both; the co-located vector split by the picture's distances
    int16_t mvps[2][2], mvs[2][2];
    mv_pred(0, mvps[0], 0);
    mv_pred(0, mvps[1], 1);
    const int tb = P.dist_tb, td = P.dist_td;
    int16_t colv[2];
    const int ncol = col_mv(colv);
    int costs[2];
    for (int l = 0; l < 2; l++) {
      int16_t extra[2][2];
      int ne = 0;
      if (ncol) {
        const int num = l == 0 ? tb : tb - td;
        extra[ne][0] = (int16_t)(colv[0] * num / td);
        extra[ne][1] = (int16_t)(colv[1] * num / td);
        ne++;
      }
      if (l == 1) {
        extra[ne][0] = (int16_t)(mvs[0][0] * (tb - td) / tb);
        extra[ne][1] = (int16_t)(mvs[0][1] * (tb - td) / tb);
        ne++;
      }
      const int cc = motion_search(0, 0, 0, 4, 4, mvps[l], (const int16_t(*)[2])extra, ne, mvs[l], l);
      costs[l] = cc + lam * (mvd_bits(mvs[l][0], mvs[l][1], mvps[l][0], mvps[l][1]) + 3);
    }
    {
      const int z = 0;
      bi_satd(1, &z, &z, 4, 4, mvs, mvs + 1);
    }
    const int bi = sat_sum(0, 0, 4, 4) + lam * (mvd_bits(mvs[0][0], mvs[0][1], mvps[0][0], mvps[0][1]) +
                                                mvd_bits(mvs[1][0], mvs[1][1], mvps[1][0], mvps[1][1]) + 5);
    sync();
    const int dir = bi < min(costs[0], costs[1]) ? 3 : costs[0] <= costs[1] ? 1 : 2;
    const int best16 = min(bi, min(costs[0], costs[1]));
    BParts cand[2];
    cand[0].part = 0;
    for (int k = 0; k < 4; k++) cand[0].dirs[k] = k == 0 ? (uint8_t)dir : 0;
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 4; k++) { cand[0].pmv[l][k][0] = mvs[l][0]; cand[0].pmv[l][k][1] = mvs[l][1]; }
    const int ncand = best_b_parts(mvs, mvps, best16, d8, cand[1]) ? 2 : 1;
    st->inter_cost = best16;
    st->inter_part = cand[ncand - 1].part;
    for (int ci = 0; ci < ncand; ci++) {
      const BParts &cd = cand[ci];
      // each 8x8 block's lists and vectors: those of the partition covering it, or the direct ones
      BMot m;
      for (int i8 = 0; i8 < 4; i8++) {
        const int k = cd.part == 0 ? 0 : cd.part == 1 ? i8 >> 1 : cd.part == 2 ? i8 & 1 : i8;
        const int d = cd.dirs[k];
        m.use[i8] = d == 0 ? dm.use[i8] : (uint8_t)d;
        for (int l = 0; l < 2; l++) {
          const bool on = (m.use[i8] >> l) & 1;
          m.mv[l][i8][0] = !on ? 0 : d == 0 ? dm.mv[l][i8][0] : cd.pmv[l][k][0];
          m.mv[l][i8][1] = !on ? 0 : d == 0 ? dm.mv[l][i8][1] : cd.pmv[l][k][1];
        }
      }
      b_pred(m);
      int ch = -1;
      code_b(cd.part, cd.dirs, m, 0);
      code_inter_residual(0);
      rd_keep(lam2, &ch, &have, &best_j);
      if (P.transform8) {
        recode_inter_t8();
        if (L.code.t8x8) rd_keep(lam2, &ch, &have, &best_j);
      }
    }
    int i16c;
    const int i16 = best_i16(&i16c);
    st->i16_cost = i16c;
    int intra_chroma = -1;
    code_i16(i16);
    code_intra_chroma();
    save_chroma();
    rd_keep(lam2, &intra_chroma, &have, &best_j);
    // Intra8x8 is not tried in B pictures (enc/frame.rs:1611)
    code_i4();
    restore_chroma();
    rd_keep(lam2, &intra_chroma, &have, &best_j);
    restore_rec(L.keep_rec);
    copy_code(&L.code, &L.keep);
    sync();
    st->rd_cost = best_j;
    rd_kept = 1;
  }

  // enc/frame.rs:715-790; the decision lands in L.code and the canvas
  __device__ void code_mb(ia_mbstats *st) {
    const int lam = P.lambda;
    rd_kept = 0;
    st->skip_sad = -1;
    st->i16_cost = -1;
    st->inter_cost = -1;
    st->inter_part = -1;
    st->rd_cost = 0.0;
    if (P.b_slice) {
      decide_b(st);
      return;
    }
    if (P.p_slice) {
      int16_t smv[2];
      skip_mv(smv);
      luma_pred(0, mbx * 16, mby * 16, 16, 16, smv[0], smv[1], L.pred, 16);
      sync();
      int s = 0;
      for (int i = lane(); i < 256; i += 64) s += iabs((int)L.src_y[i] - (int)L.pred[i]);
      const int sad = wave_sum(s);
      sync();
      st->skip_sad = sad;
      if (P.fast && (double)sad < P.fast_skip_y) {
        // realtime: P_Skip without a transform when luma and chroma are within a twentieth of a
        // quantiser step of the skip prediction (enc/frame.rs:731-746)
        int16_t mvs[16][2];
        for (int i = 0; i < 16; i++) { mvs[i][0] = smv[0]; mvs[i][1] = smv[1]; }
        const uint8_t z4[4] = {0, 0, 0, 0};
        inter_pred(0, z4, mvs, z4);
        sync();
        int cs = 0;
        for (int i = lane(); i < 128; i += 64) cs += iabs((int)L.src_c[i >> 6][i & 63] - (int)L.cpred[i >> 6][i & 63]);
        const int csad = wave_sum(cs);
        if ((double)csad < P.fast_skip_c) {
          code_new(K_INTER);
          if (lane() == 0) {
            L.code.skip = 1;
            for (int i = 0; i < 16; i++) { L.code.mvs[i][0] = smv[0]; L.code.mvs[i][1] = smv[1]; }
          }
          for (int i = lane(); i < 256; i += 64) cyv(i & 15, i >> 4) = L.pred[i];
          for (int i = lane(); i < 128; i += 64) ccv(i >> 6, (i & 63) & 7, (i & 63) >> 3) = L.cpred[i >> 6][i & 63];
          sync();
          return;
        }
        sync();
      }
      if ((double)sad < P.skip_thresh) {
        int16_t mvs[16][2];
        for (int i = 0; i < 16; i++) { mvs[i][0] = smv[0]; mvs[i][1] = smv[1]; }
        const uint8_t z4[4] = {0, 0, 0, 0};
        code_inter(0, mvs, z4, z4, 0);
        if (L.code.cbp == 0) {
          if (lane() == 0) L.code.skip = 1;
          sync();
          return;
        }
      }
    }
    int i16c;
    const int i16 = best_i16(&i16c);
    st->i16_cost = i16c;
    int best_intra = i16c + lam * 4;
    Inter inter;
    if (P.p_slice) {
      best_inter(inter);
      st->inter_cost = inter.cost;
      st->inter_part = inter.part;
    }
    if (P.use_cabac) {
      decide_rd(i16, P.p_slice ? &inter : nullptr, st);
      return;
    }
    // SATD decisions (CAVLC), enc/frame.rs:760-789
    // realtime: flat intra macroblocks stay Intra16x16 (enc/frame.rs:763-765)
    const int try_i4 = (!P.p_slice || best_intra < inter.cost) && !(P.fast && !P.p_slice && (double)i16c < P.fast_flat);
    int have_i4 = 0;
    if (try_i4) {
      const int cost = code_i4() + lam * 6;
      if (cost < best_intra) {
        best_intra = cost;
        have_i4 = 1;
        copy_code(&L.i4, &L.code);
        save_rec(L.keep_rec);
        sync();
      }
    }
    if (P.p_slice && inter.cost <= best_intra) {
      code_inter(inter.part, inter.mvs, inter.refs, inter.sub, 0);
      int16_t sm[2];
      skip_mv(sm);
      if (L.code.cbp == 0 && inter.part == 0 && inter.mvs[0][0] == sm[0] && inter.mvs[0][1] == sm[1] && inter.refs[0] == 0) {
        if (lane() == 0) L.code.skip = 1;
        sync();
      }
      return;
    }
    if (have_i4) {
      copy_code(&L.code, &L.i4);
      restore_rec(L.keep_rec);
      sync();
    } else {
      code_i16(i16);
    }
    code_intra_chroma();
  }

  // enc/frame.rs:2281-2339: the macroblock's state for later macroblocks and its MbInfo
  __device__ void finish_mb(const ia_mbcode &code, MbState &s, ia_mbinfo &info) const {
    const int inter = code.kind == K_INTER, bm = inter && P.b_slice;
    const uint8_t z4[4] = {0, 0, 0, 0};
    // the lists each 8x8 block uses
    int used0[4], used1[4];
    for (int b8 = 0; b8 < 4; b8++) {
      used0[b8] = inter && (!bm || (code.buse[b8] & 1));
      used1[b8] = bm && ((code.buse[b8] >> 1) & 1);
    }
    Rect rs[16];
    const int n = rects(code.part, bm ? z4 : code.sub, rs);
    s = MbState{};
    if (inter && !code.skip) {
      if (bm) {
        mvd_grid(rs, n, code.bmvd[0], s.mvd);
        mvd_grid(rs, n, code.bmvd[1], s.mvd1);
      } else {
        mvd_grid(rs, n, code.mvds, s.mvd);
      }
    }
    if (bm) {
      for (int b = 0; b < 16; b++) {
        const int b8 = (b / 8) * 2 + (b % 4) / 2;
        for (int q = 0; q < 2; q++) {
          s.mvs[b][q] = used0[b8] ? code.bmv[0][b][q] : 0;
          s.mvs1[b][q] = used1[b8] ? code.bmv[1][b][q] : 0;
        }
      }
    } else if (inter) {
      for (int k = 0; k < n; k++)
        for (int by = rs[k].y; by < rs[k].y + rs[k].h; by++)
          for (int bx = rs[k].x; bx < rs[k].x + rs[k].w; bx++) {
            s.mvs[by * 4 + bx][0] = code.mvs[k][0];
            s.mvs[by * 4 + bx][1] = code.mvs[k][1];
          }
    }
    s.bdirect = (uint8_t)(bm && code.part == 0 && code.bdir[0] == 0);
    const int cbpc = code.cbp >> 4;
    s.cbf_dc = (uint8_t)((code.kind == K_I16 && count_nz(code.dc, 16) > 0) | ((cbpc > 0 && count_nz(code.cdc[0], 4) > 0) << 1) |
                         ((cbpc > 0 && count_nz(code.cdc[1], 4) > 0) << 2));
    int8_t r16[16];
    for (int k = 0; k < NPARTS[code.part]; k++) {
      const uint8_t *pr = PARTS[code.part][k];
      for (int by = pr[1]; by < pr[1] + pr[3]; by++)
        for (int bx = pr[0]; bx < pr[0] + pr[2]; bx++) r16[by * 4 + bx] = (int8_t)code.refs[k];
    }
    int8_t r8[4];
    for (int b8 = 0; b8 < 4; b8++) r8[b8] = r16[(b8 / 2) * 8 + (b8 % 2) * 2];
    s.i4 = code.kind == K_I4 || code.kind == K_I8;
    s.t8x8 = code.t8x8;
    s.inter = (uint8_t)inter;
    for (int i = 0; i < 16; i++) {
      s.modes[i] = code.modes[i];
      s.nnz[i] = code.nnz[i];
    }
    for (int c = 0; c < 2; c++)
      for (int k = 0; k < 4; k++) s.cnnz[c][k] = code.cnnz[c][k];
    for (int b8 = 0; b8 < 4; b8++) {
      s.refs[b8] = (int8_t)(inter && !bm ? r8[b8] : used0[b8] ? 0 : -1);
      s.refs1[b8] = (int8_t)(used1[b8] ? 0 : -1);
    }
    s.skip = code.skip;
    s.kind = code.kind;
    s.cbp = code.cbp;
    s.chroma_mode = inter ? 0 : code.chroma_mode;
    info = ia_mbinfo{};
    info.slice = (uint16_t)slice_id;
    info.kind = code.kind;
    info.qp = (int8_t)P.qp_slice;
    for (int i = 0; i < 16; i++) info.nnz[i] = code.nnz[i];
    info.cbp = code.cbp;
    info.skip = code.skip;
    info.t8x8 = code.t8x8;
    for (int i = 0; i < 16; i++)
      for (int q = 0; q < 2; q++) {
        info.mv[i][q] = s.mvs[i][q];
        info.mv1[i][q] = s.mvs1[i][q];
      }
    for (int b8 = 0; b8 < 4; b8++) {
      // the identity tells the loop filter which picture a block predicts from (B: one per list)
      const int r = bm ? 0 : r8[b8];
      info.ref_idx[b8] = (int8_t)(used0[b8] ? r : -1);
      info.ref_id[b8] = used0[b8] ? (bm ? 1 : 1 + r) : -1;
      info.ref_idx1[b8] = (int8_t)(used1[b8] ? 0 : -1);
      info.ref_id1[b8] = used1[b8] ? 2 : -1;
    }
    info.bdirect = s.bdirect;
    if (s.bdirect) info.direct8 = 15;
    else if (bm && code.part == 3)
      for (int k = 0; k < 4; k++)
        if (code.bdir[k] == 0) info.direct8 |= (uint8_t)(1 << k);
  }

  // ---- the slice loop (enc/frame.rs:452-487) ----

  __device__ void load_mb() {
    const int ls_ = ls(), cs_ = cs();
    const int x0 = mbx * 16, y0 = mby * 16;
    for (int i = lane(); i < 256; i += 64) L.src_y[i] = P.src[0][(size_t)(y0 + (i >> 4)) * ls_ + x0 + (i & 15)];
    for (int i = lane(); i < 128; i += 64) {
      const int ci = i >> 6, o = i & 63;
      L.src_c[ci][o] = P.src[1 + ci][(size_t)(mby * 8 + (o >> 3)) * cs_ + mbx * 8 + (o & 7)];
    }
    // canvas edges from the reconstruction of earlier macroblocks of the slice
    const int ta = up(), la = mbx > 0;
    for (int i = lane(); i < 28; i += 64) {
      const int x = i - 1;  // -1..26
      uint8_t v = 0;
      if (ta && x0 + x >= 0 && x0 + x < ls_ && x <= 23 && (x >= 0 || la)) v = P.rec[0][(size_t)(y0 - 1) * ls_ + x0 + x];
      L.cy[0][i] = v;
    }
    for (int i = lane(); i < 16; i += 64) L.cy[i + 1][0] = la ? P.rec[0][(size_t)(y0 + i) * ls_ + x0 - 1] : 0;
    for (int i = lane(); i < 2 * 9; i += 64) {
      const int ci = i / 9, x = i % 9 - 1;
      uint8_t v = 0;
      if (ta && (x >= 0 || la)) v = P.rec[1 + ci][(size_t)(mby * 8 - 1) * cs_ + mbx * 8 + x];
      L.cc[ci][0][x + 1] = v;
    }
    for (int i = lane(); i < 16; i += 64) {
      const int ci = i >> 3, y = i & 7;
      L.cc[ci][y + 1][0] = la ? P.rec[1 + ci][(size_t)(mby * 8 + y) * cs_ + mbx * 8 - 1] : 0;
    }
    // neighbour states: left, top, top-right, top-left
    const int nx[4] = {mbx - 1, mbx, mbx + 1, mbx - 1}, ny[4] = {mby, mby - 1, mby - 1, mby - 1};
    for (int k = 0; k < 4; k++) {
      const int ok = nx[k] >= 0 && nx[k] < P.mb_w && ny[k] >= row0 && ny[k] < P.mb_h;
      if (lane() == 0) {
        L.nb_ok[k] = ok;
        if (ok) L.nb[k] = P.state[ny[k] * P.mb_w + nx[k]];
      }
    }
    sync();
  }

  __device__ void store_mb(int idx, const ia_mbstats &st) {
    const int ls_ = ls(), cs_ = cs();
    const int x0 = mbx * 16, y0 = mby * 16;
    for (int i = lane(); i < 256; i += 64) P.rec[0][(size_t)(y0 + (i >> 4)) * ls_ + x0 + (i & 15)] = cyv(i & 15, i >> 4);
    for (int i = lane(); i < 128; i += 64) {
      const int ci = i >> 6, o = i & 63;
      P.rec[1 + ci][(size_t)(mby * 8 + (o >> 3)) * cs_ + mbx * 8 + (o & 7)] = ccv(ci, o & 7, o >> 3);
    }
    // the code as it will be written
    const uint32_t *s = reinterpret_cast<const uint32_t *>(&L.code);
    uint32_t *d = reinterpret_cast<uint32_t *>(&P.codes[idx]);
    for (int i = lane(); i < (int)(sizeof(ia_mbcode) / 4); i += 64) d[i] = s[i];
    // the chosen macroblock's bins adapt the live contexts as the writer would; their count is its rate
    uint32_t bits = 0;
    if (P.use_cabac) {
      if (rd_kept) {
        // counted when it was costed: take its rate and final states
        const uint32_t *s = reinterpret_cast<const uint32_t *>(L.kept_st);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.cabac);
        for (int i = lane(); i < 110; i += 64) d[i] = s[i];
        bits = L.kept_bits;
      } else {
        int chroma = -1;
        bits = count_bins(L.cabac, &chroma);
      }
    }
    // the state, info and statistics formed in LDS by one lane, then stored a dword per lane (the
    // three structures are whole dwords)
    if (lane() == 0) {
      finish_mb(L.code, L.fin.st, L.fin.info);
      L.fin.stats = st;
      L.fin.stats.bits = (int32_t)bits;
    }
    sync();
    auto store_words = [&](void *dst, const void *src, int bytes) {
      const uint32_t *s4 = reinterpret_cast<const uint32_t *>(src);
      uint32_t *d4 = reinterpret_cast<uint32_t *>(dst);
      for (int i = lane(); i < bytes / 4; i += 64) d4[i] = s4[i];
    };
    static_assert(sizeof(MbState) % 4 == 0 && sizeof(ia_mbinfo) % 4 == 0 && sizeof(ia_mbstats) % 4 == 0, "whole dwords");
    store_words(&P.state[idx], &L.fin.st, (int)sizeof(MbState));
    store_words(&P.mbs[idx], &L.fin.info, (int)sizeof(ia_mbinfo));
    store_words(&P.stats[idx], &L.fin.stats, (int)sizeof(ia_mbstats));
    __threadfence_block();
    sync();
  }
};

// Slice s of the picture p describes, its macroblocks in raster order.
__device__ void analyse_slice(const AnParams &p, int s) {
  Coder c(p);
  c.row0 = s * p.mb_h / p.nslices;
  c.row1 = (s + 1) * p.mb_h / p.nslices;
  c.slice_id = s + 1;
  // context states of the slice (enc/cabac.rs:54-64)
  if (p.use_cabac) {
    const int qp = clampi(p.qp_slice, 0, 51);
    for (int i = threadIdx.x; i < 1024; i += 64) {
      const int m = p.p_slice ? IA_CABAC_INIT_P[i][0] : IA_CABAC_INIT_I[i][0];
      const int n = p.p_slice ? IA_CABAC_INIT_P[i][1] : IA_CABAC_INIT_I[i][1];
      const int pre = clampi(((m * qp) >> 4) + n, 1, 126);
      L.cabac[i] = pre <= 63 ? (uint8_t)((63 - pre) << 1) : (uint8_t)(((pre - 64) << 1) | 1);
    }
    // the 8x8 scan positions (0-62) that code on each significance and each last context
    if (threadIdx.x < 24) {
      const int j = threadIdx.x;
      uint64_t m = 0;
      for (int q = 0; q < 63; q++)
        if ((j < 15 ? IA_SIG_OFF_8X8[0][q] : IA_LAST_OFF_8X8[q] + 15) == j) m |= 1ull << q;
      L.pos8[j] = m;
    }
    // a decision's cost and next state per (state, LPS?): state s = pStateIdx << 1 | valMPS
    for (int i = threadIdx.x; i < 256; i += 64) {
      const int st = i >> 1, lps = i & 1, ps = st >> 1, mps = st & 1;
      const uint32_t ns = lps ? (uint32_t)((IA_TRANS_LPS[ps] << 1) | (ps == 0 ? !mps : mps))
                              : (uint32_t)(((ps < 62 ? ps + 1 : ps) << 1) | mps);
      L.cab[i] = (p.entropy[ps * 2 + lps] << 8) | ns;
    }
  }
  __syncthreads();
  for (int mby = c.row0; mby < c.row1; mby++)
    for (int mbx = 0; mbx < p.mb_w; mbx++) {
      c.mbx = mbx;
      c.mby = mby;
      c.load_mb();
      ia_mbstats st;
      c.code_mb(&st);
      c.store_mb(mby * p.mb_w + mbx, st);
    }
}

// Three waves per SIMD: without an occupancy target the compiler spreads the analysis over 391
// registers (AGPRs as spill space), one wave per SIMD; held to 168 it adds no register spills (its
// private segment grows by 64 bytes per lane). The kernel is latency-bound, so the added waves are
// nearly free throughput: 1.9x on 64 streams of mixed sizes.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void analyse_kernel(AnParams p) {
  analyse_slice(p, blockIdx.x);
}

// Many pictures at once: a pool of workgroups takes (picture, slice) tasks from a queue, tasks[t] =
// picture << 16 | slice for t < ntasks (longest first), ps[picture] its parameters. *head counts the
// tasks taken and must be 0 at the launch; a worker leaves when the queue is empty.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void analyse_batch_kernel(const AnParams *ps,
                                                                                                    const int *tasks,
                                                                                                    int ntasks, int *head) {
  __shared__ int next;
  for (;;) {
    __syncthreads();  // the previous task is done with `next` and the LDS
    if (threadIdx.x == 0) next = atomicAdd(head, 1);
    __syncthreads();
    const int t = next;
    if (t >= ntasks) return;
    const int e = __builtin_amdgcn_readfirstlane(tasks[t]);
    analyse_slice(ps[e >> 16], e & 0xffff);
  }
}

}  // namespace

hipError_t launch_analyse_batch(const AnParams *ps, const int *tasks, int ntasks, int *head, int workers,
                                hipStream_t stream) {
  hipLaunchKernelGGL(analyse_batch_kernel, dim3(workers), dim3(64), 0, stream, ps, tasks, ntasks, head);
  return hipGetLastError();
}

int analyse_batch_workers() {
  int dev = 0, per = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, analyse_batch_kernel, 64, 0) != hipSuccess) return 0;
  return (per > 0 ? per : 1) * prop.multiProcessorCount;
}

hipError_t launch_analyse(const AnParams &p, hipStream_t stream) {
  hipLaunchKernelGGL(analyse_kernel, dim3(p.nslices), dim3(64), 0, stream, p);
  return hipGetLastError();
}

}  // namespace ia
