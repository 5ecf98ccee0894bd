// halfpel.hip — the half-sample planes of a reference picture on gfx950.
//
// Builds, from one 8-bit luma plane, the planes `HalfPel::build` builds
// (crates/ironav-codec/src/h264/enc/subpel.rs:57-115):
//   G  the integer samples with the picture's edges repeated PAD samples outwards,
//   B  the horizontal 6-tap half samples (v+16)>>5, columns 2..sw-4 (0 elsewhere),
//   H  the vertical 6-tap half samples, rows 2..sh-4 (0 elsewhere),
//   J  the vertical 6-tap over the unrounded horizontal sums, (v+512)>>10, rows 2..sh-4; the sums are
//      0 outside columns 2..sw-4 as in the reference's scratch buffer.
// HBM-bound (4 output bytes per sample written once, the luma read about once). One wave builds a
// strip of HP_TW = 256 columns x HP_TH rows of all four planes: a lane owns 4 adjacent columns and
// walks down the strip, keeping the last six rows' G samples and horizontal sums in registers, so
// the vertical taps need no LDS and no barrier. Per row a lane reads three aligned luma dwords (its
// own 4 columns and the 2/3-column reach either side; neighbouring lanes share the cache lines) and
// writes one dword per plane; a wave's stores cover 256 contiguous bytes of a plane row. All of a
// strip's loads are issued before its first use; a single picture's waves each build several strips
// of one column band and fetch the next strip's rows before building the current one.
#include <algorithm>

#include "common.h"

namespace ia {
namespace {

constexpr int TW = HP_TW, TH = HP_TH;
static_assert(TW == 256, "a wave of 64 lanes x 4 columns");

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const uint32_t GU32;
typedef __attribute__((address_space(1))) uint32_t G32;

__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// 6-tap filter (1, -5, 20, 20, -5, 1) on two 16-bit lanes at once. The sums of 8-bit samples lie in
// [-2550, 10710], so the lanes' wrapping arithmetic gives them exactly.
__device__ __forceinline__ s16x2 tap6x2(s16x2 a, s16x2 b, s16x2 c, s16x2 d, s16x2 e, s16x2 f) {
  return (a + f) - (b + e) * (s16x2){5, 5} + (c + d) * (s16x2){20, 20};
}

// clip(0, 255, (v + 16) >> 5) on both lanes; the clip is applied before the shift (the same: the
// upper bound is 256 * 32 - 1), leaving each result in the low byte of its lane.
__device__ __forceinline__ uint32_t round_clip5x2(s16x2 v) {
  const s16x2 t = __builtin_elementwise_min(__builtin_elementwise_max(v + (s16x2){16, 16}, (s16x2){0, 0}),
                                            (s16x2){8191, 8191});
  return as_u32(t >> (s16x2){5, 5});
}

// The four bytes 0 and 2 of `lo` and of `hi` (two pairs of 8-bit results in 16-bit lanes), in order.
__device__ __forceinline__ uint32_t pack_bytes(uint32_t lo, uint32_t hi) {
  return __builtin_amdgcn_perm(hi, lo, 0x06040200u);
}

// One lane's place in a strip: its columns, luma dword offsets and border handling.
//
// Per G row the lane forms its samples as 16-bit pairs of adjacent columns (`v_perm_b32` from the
// loaded dwords) and filters two columns per packed instruction: B's and H's sums fit 16 bits. J's
// sum (up to 42 x 10710) does not: it is a dot product over the six rows' horizontal sums, taken
// as three `v_dot2_i32_i16` over pairs (row r-1, row r) of one column, with 32-bit accumulation.
struct Lane {
  int x;            // first of its four columns (a multiple of 4)
  int off[3];       // dword index in a luma row of G columns x-4.., x.., x+4..
  uint32_t sel[3];  // `v_perm_b32` selector: the dword itself, or its first / last byte repeated
  uint32_t cmask[2];  // 16-bit lane masks of the column pairs whose horizontal sums are kept
  bool edge;          // a border dword or a dropped horizontal sum among its columns
};

// G columns x-4.., x.., x+4.. are luma columns lc-4, lc, lc+4 (multiples of 4, as pw is). A dword
// lies wholly inside the picture or wholly in the left or right border, where G repeats the row's
// first or last sample: the lane loads the nearest dword inside and, in a border, spreads its first
// or last byte over all four. The horizontal sums are 0 outside columns 2..sw-4.
__device__ __forceinline__ Lane lane_setup(int x, int pw) {
  const int sw = pw + 2 * PAD, lc = x - PAD;
  Lane L;
  L.x = x;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int c = lc + 4 * (q - 1);
    L.off[q] = min(max(c, 0), pw - 4) >> 2;
    L.sel[q] = c < 0 ? 0x00000000u : c >= pw ? 0x03030303u : 0x03020100u;
  }
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int c0 = x + 2 * p;
    L.cmask[p] = (c0 >= 2 && c0 <= sw - 4 ? 0x0000ffffu : 0u) | (c0 + 1 >= 2 && c0 + 1 <= sw - 4 ? 0xffff0000u : 0u);
  }
  L.edge = L.sel[0] != 0x03020100u || L.sel[2] != 0x03020100u || (L.cmask[0] & L.cmask[1]) != 0xffffffffu;
  return L;
}

// The lane's three dwords of G rows y0-2 .. y0+TH+2.
__device__ __forceinline__ void load_rows(uint32_t (&d)[TH + 5][3], const Lane &L, const uint8_t *__restrict__ luma,
                                          int pw, int ph, int y0) {
#pragma unroll
  for (int i = 0; i < TH + 5; i++) {
    const int lr = min(max(y0 - 2 + i - PAD, 0), ph - 1);
    GU32 *row = (GU32 *)(luma + (size_t)lr * pw);
#pragma unroll
    for (int q = 0; q < 3; q++) d[i][q] = row[L.off[q]];
  }
}

// Rows y0 .. y0+TH-1 (those inside the planes) of the lane's four columns of the four planes, from
// the G rows `load_rows` fetched. EDGE: the wave has lanes at the planes' left or right border.
template <bool EDGE>
__device__ __forceinline__ void build_rows(const uint32_t (&d)[TH + 5][3], const Lane &L, int pw, int ph,
                                           uint8_t *__restrict__ planes, int y0) {
  const int sw = pw + 2 * PAD, sh = ph + 2 * PAD;
  const size_t n4 = (size_t)sw * sh / 4;  // dwords per plane (sw is a multiple of 16)
  const int x = L.x;
  const uint32_t *sel = L.sel, *cmask = L.cmask;
  // rings over the last six rows, indexed by row % 6
  s16x2 gp[6][2];   // G of columns (0,1) and (2,3)
  s16x2 bp[6][2];   // unrounded horizontal sums of the same column pairs
  s16x2 rp[6][4];   // per column, its horizontal sums of rows (r-1, r)
#pragma unroll
  for (int i = 0; i < TH + 5; i++) {
    const uint32_t w0 = EDGE ? __builtin_amdgcn_perm(d[i][0], d[i][0], sel[0]) : d[i][0];
    const uint32_t w1 = EDGE ? __builtin_amdgcn_perm(d[i][1], d[i][1], sel[1]) : d[i][1];
    const uint32_t w2 = EDGE ? __builtin_amdgcn_perm(d[i][2], d[i][2], sel[2]) : d[i][2];
    // (G(x+a), G(x+a+1)) as 16-bit lanes, a = -2 .. 5
    const s16x2 pm2 = as_s16x2(__builtin_amdgcn_perm(w0, w0, 0x0c030c02u));
    const s16x2 pm1 = as_s16x2(__builtin_amdgcn_perm(w0, w1, 0x0c000c07u));
    const s16x2 p0 = as_s16x2(__builtin_amdgcn_perm(w1, w1, 0x0c010c00u));
    const s16x2 p1 = as_s16x2(__builtin_amdgcn_perm(w1, w1, 0x0c020c01u));
    const s16x2 p2 = as_s16x2(__builtin_amdgcn_perm(w1, w1, 0x0c030c02u));
    const s16x2 p3 = as_s16x2(__builtin_amdgcn_perm(w1, w2, 0x0c000c07u));
    const s16x2 p4 = as_s16x2(__builtin_amdgcn_perm(w2, w2, 0x0c010c00u));
    const s16x2 p5 = as_s16x2(__builtin_amdgcn_perm(w2, w2, 0x0c020c01u));
    const int r = i % 6;
    gp[r][0] = p0;
    gp[r][1] = p2;
    bp[r][0] = tap6x2(pm2, pm1, p0, p1, p2, p3);
    bp[r][1] = tap6x2(p0, p1, p2, p3, p4, p5);
    if (EDGE) {
      bp[r][0] = as_s16x2(as_u32(bp[r][0]) & cmask[0]);
      bp[r][1] = as_s16x2(as_u32(bp[r][1]) & cmask[1]);
    }
    if (i >= 1) {
      const int rm = (i - 1) % 6;
#pragma unroll
      for (int k = 0; k < 4; k++)
        rp[r][k] = as_s16x2(__builtin_amdgcn_perm(as_u32(bp[r][k >> 1]), as_u32(bp[rm][k >> 1]),
                                                  (k & 1) ? 0x07060302u : 0x05040100u));
    }
    if (i < 5) continue;
    // rows y-2 .. y+3 of output row y = y0 + i - 5 are in the rings
    const int y = y0 + i - 5;
    if (y >= sh) break;
    const bool vrow = y >= 2 && y < sh - 3;
    const int c = (i - 3) % 6, r0 = (i - 5) % 6, r1 = (i - 4) % 6, r2 = c, r3 = (i - 2) % 6, r4 = (i - 1) % 6;
    const uint32_t ob = pack_bytes(round_clip5x2(bp[c][0]), round_clip5x2(bp[c][1]));
    uint32_t oh = pack_bytes(round_clip5x2(tap6x2(gp[r0][0], gp[r1][0], gp[r2][0], gp[r3][0], gp[r4][0], gp[r][0])),
                             round_clip5x2(tap6x2(gp[r0][1], gp[r1][1], gp[r2][1], gp[r3][1], gp[r4][1], gp[r][1])));
    uint32_t oj = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int acc = __builtin_amdgcn_sdot2(rp[r1][k], (s16x2){1, -5}, 512, false);
      acc = __builtin_amdgcn_sdot2(rp[r3][k], (s16x2){20, 20}, acc, false);
      acc = __builtin_amdgcn_sdot2(rp[r][k], (s16x2){-5, 1}, acc, false);
      oj |= ((uint32_t)min(max(acc, 0), 262143) >> 10) << (8 * k);
    }
    if (!vrow) oh = oj = 0;
    // row y of plane G (uniform address) and the lane's dword in it and in the three planes after
    G32 *prow = (G32 *)(planes + (size_t)y * sw);
    const int q = x >> 2;
    prow[q] = EDGE ? __builtin_amdgcn_perm(d[i - 3][1], d[i - 3][1], sel[1]) : d[i - 3][1];
    prow[n4 + q] = ob;
    prow[2 * n4 + q] = oh;
    prow[3 * n4 + q] = oj;
  }
}

// The border handling only where the wave has border lanes (a wave-uniform branch).
__device__ __forceinline__ void build_strip(const uint32_t (&d)[TH + 5][3], const Lane &L, int pw, int ph,
                                            uint8_t *__restrict__ planes, int y0) {
  if (__ballot(L.edge)) build_rows<true>(d, L, pw, ph, planes, y0);
  else build_rows<false>(d, L, pw, ph, planes, y0);
}

// One picture: wave w builds the strips of column band w % tiles_x, row bands w / tiles_x, + rows,
// ... (rows = gridDim.x / tiles_x), fetching the next strip's G rows before building the current.
__global__ __launch_bounds__(64) void halfpel_kernel(const uint8_t *__restrict__ luma, int pw, int ph,
                                                      uint8_t *__restrict__ planes, int tiles_x, int tiles_y) {
  const int bands = gridDim.x / tiles_x;
  const int bx = blockIdx.x % tiles_x;
  int by = blockIdx.x / tiles_x;
  const int x = bx * TW + 4 * (int)threadIdx.x;
  if (x >= pw + 2 * PAD || by >= tiles_y) return;  // sw is a multiple of 4: all four columns or none
  const Lane L = lane_setup(x, pw);
  uint32_t d[TH + 5][3];
  load_rows(d, L, luma, pw, ph, by * TH);
  for (;;) {
    const int nb = by + bands;
    uint32_t e[TH + 5][3];
    if (nb < tiles_y) load_rows(e, L, luma, pw, ph, nb * TH);
    build_strip(d, L, pw, ph, planes, by * TH);
    if (nb >= tiles_y) break;
    by = nb;
#pragma unroll
    for (int i = 0; i < TH + 5; i++)
#pragma unroll
      for (int q = 0; q < 3; q++) d[i][q] = e[i][q];
  }
}

// Many pictures: block b builds strip b - first of the job whose strips hold it (jobs by `first`).
__global__ __launch_bounds__(64) void halfpel_batch_kernel(const HpJob *jobs, int njobs) {
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= b) lo = mid;
    else hi = mid - 1;
  }
  const HpJob &j = jobs[lo];
  if (!j.planes) return;
  const int t = b - j.first;
  const int x = (t % j.tiles_x) * TW + 4 * (int)threadIdx.x, y0 = (t / j.tiles_x) * TH;
  if (x >= j.pw + 2 * PAD) return;
  const Lane L = lane_setup(x, j.pw);
  uint32_t d[TH + 5][3];
  load_rows(d, L, j.luma, j.pw, j.ph, y0);
  build_strip(d, L, j.pw, j.ph, j.planes, y0);
}

}  // namespace

hipError_t launch_halfpel_batch(const HpJob *jobs, int njobs, int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL(halfpel_batch_kernel, dim3(nblocks), dim3(64), 0, stream, jobs, njobs);
  return hipGetLastError();
}

hipError_t launch_halfpel(const uint8_t *luma, int pw, int ph, uint8_t *planes, hipStream_t stream) {
  if ((reinterpret_cast<uintptr_t>(luma) & 3) || (pw & 3) || pw < 4 || ph < 1) return hipErrorInvalidValue;
  const int sw = pw + 2 * PAD, sh = ph + 2 * PAD;
  const int tiles_x = (sw + TW - 1) / TW, tiles_y = (sh + TH - 1) / TH;
  const int bands = std::min(tiles_y, std::max(1, HP_WAVES / tiles_x));
  hipLaunchKernelGGL(halfpel_kernel, dim3(tiles_x * bands), dim3(64), 0, stream, luma, pw, ph, planes, tiles_x, tiles_y);
  return hipGetLastError();
}

}  // namespace ia
