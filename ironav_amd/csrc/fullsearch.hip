// fullsearch.hip — exhaustive integer-sample SAD search of all 41 partitions on gfx950.
//
// One workgroup searches one macroblock against one reference over the window of ia_motion_search
// (include/ironav_h264a.h): every integer displacement (dx, dy) the whole macroblock may take
// (crates/ironav-codec/src/h264/enc/frame.rs:952-954) within +-range. It returns, per partition, the
// least SAD and its displacement, ties to the first in raster order of (dy, dx).
//
// Design (integer VALU bound; SAD is not a dense contraction, so no MFMA):
//  - The reference window is staged once in LDS (coalesced loads, clipped at the plane's edges). For
//    ranges up to 32 every byte alignment of a row is stored once, so a candidate's 16 samples of a
//    row are 4 dwords read as they are; for ranges up to 64 (a 4x larger window) rows are aligned
//    dwords and a candidate's row is 5 of them realigned by v_alignbyte.
//  - A lane owns one column of displacements and walks a strip of K rows of it: each reference row
//    loaded from LDS feeds the K candidates that use it (one row is read once per K candidates).
//  - The 64 source dwords of the macroblock are uniform: scalar loads put them in scalar registers.
//  - v_sad_hi_u8 accumulates each 4x4 block's SAD into the high half of a 32-bit word whose low half
//    holds the candidate's row index in the window. Partition SADs are sums of 4x4 words, so every
//    partition's word is already a sort key (SAD, then row) at one add per partition, and a running
//    minimum per partition costs one v_min3 per two candidates.
//  - A lane's candidates are taken in increasing row order, and within one row (possible when the
//    window is wider than the workgroup) in increasing column order, so the low half orders them as
//    row * TQ + column step, TQ being the most columns of one row a lane visits. At the end the
//    wave's least key per partition is found by DPP and decoded to (SAD, raster index) once, and the
//    waves' results meet in LDS.
#include "common.h"

namespace ia {
namespace {

constexpr int RMAX = 64;
constexpr int NXCD = 8;  // MI355X: accelerator complex dies, each with its own L2

__device__ __forceinline__ uint32_t ld_u32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const uint32_t CU32;  // constant address space: scalar loads

// v_min3_u32 written out: the compiler otherwise builds a tree over the candidates of a strip and
// merges it into the running minimum, one instruction more per partition and strip.
__device__ __forceinline__ uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_min3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// The 41 partition words of one candidate from its 16 4x4 words (raster order), numbered as
// IA_NPART documents. A word is (SAD << 16) + (4x4 blocks summed) * row, so one add per partition.
__device__ __forceinline__ void partition_keys(const uint32_t *b, uint32_t *key) {
#pragma unroll
  for (int b8 = 0; b8 < 4; b8++) {
    const int o = (b8 >> 1) * 8 + (b8 & 1) * 2;  // raster index of the 8x8 block's first 4x4
    const uint32_t t = b[o] + b[o + 1], bt = b[o + 4] + b[o + 5];
    key[9 + 2 * b8] = t;
    key[10 + 2 * b8] = bt;
    key[17 + 2 * b8] = b[o] + b[o + 4];
    key[18 + 2 * b8] = b[o + 1] + b[o + 5];
    key[5 + b8] = t + bt;
    key[25 + 4 * b8] = b[o];
    key[26 + 4 * b8] = b[o + 1];
    key[27 + 4 * b8] = b[o + 4];
    key[28 + 4 * b8] = b[o + 5];
  }
  key[1] = key[5] + key[6];
  key[2] = key[7] + key[8];
  key[3] = key[5] + key[7];
  key[4] = key[6] + key[8];
  key[0] = key[1] + key[2];
}

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v) {
  // lanes whose DPP source is out of range keep their value (bound_ctrl off, old = identity)
  return min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)v, CTRL, ROW_MASK, BANK_MASK, false));
}

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ void dpp_min_all(uint32_t *m) {
#pragma unroll
  for (int p = 0; p < IA_NPART; p++) m[p] = dpp_min<CTRL, ROW_MASK, BANK_MASK>(m[p]);
}

// Minimum over the 64 lanes of a wave: an inclusive scan by DPP (row shifts of the running minimum by
// 1, 2, 4 and 8 lanes, then row broadcasts), whose last lane holds the whole wave's minimum.
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  v = dpp_min<0x111, 0xf, 0xf>(v);
  v = dpp_min<0x112, 0xf, 0xf>(v);
  v = dpp_min<0x114, 0xf, 0xe>(v);
  v = dpp_min<0x118, 0xf, 0xc>(v);
  v = dpp_min<0x142, 0xa, 0xf>(v);
  v = dpp_min<0x143, 0xc, 0xf>(v);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Two staged layouts, by the largest range a kernel serves:
//  - ranges up to 32, "shifted": staged row dword e holds the plane's 4 bytes from byte al + e on (al
//    the window's first dword boundary), so every byte alignment is stored once and a candidate's 4
//    dwords of a row are e = b0, b0 + 4, b0 + 8, b0 + 12: no realignment in the loop, and 64 lanes
//    at consecutive columns read 64 consecutive dwords. The last candidate column starts at byte
//    b0 <= 3 + 64, so a row holds 80 dwords (20 per alignment).
//  - ranges up to 64: aligned dwords (a row of 4x fewer dwords keeps the larger window within LDS);
//    the loop realigns 5 dwords into 4 with v_alignbyte.
constexpr int row_dwords(int rmax) { return rmax == 32 ? 4 * (((3 + 2 * rmax) >> 2) + 4) : ((3 + 2 * rmax + 16) >> 2) + 2; }
constexpr int RW32 = row_dwords(32), RW64 = row_dwords(64);
constexpr int rmax_of(int rw) { return rw == RW32 ? 32 : 64; }
// Staged rows a K-row strip walk reads: strips start at rows 0, K, ... below the window's last row
// 2 * range and read K + 15 rows each.
inline int rows_for(int k, int range) { return (2 * range + k) / k * k + 15; }

// K: candidates per strip (rows of one column); NT: threads; RAWS: dwords per staged window row
// (row_dwords of the layout); WPE: waves per SIMD the register allocation must allow.
template <int K, int NT, int RAWS>
__device__ __forceinline__ void search_one(const FsArgs &a, int rows_lds, int mb, int r, uint32_t *raw) {
  const int mbx = mb % a.mb_w, mby = mb / a.mb_w;
  const int pw = a.mb_w * 16, ph = a.mb_h * 16, sw = pw + 2 * PAD, sh = ph + 2 * PAD;
  const int x0 = mbx * 16, y0 = mby * 16, R = a.range;
  const int dxlo = max(-R, -(x0 + 24)), dxhi = min(R, pw - x0 + 8);
  const int dylo = max(-R, -(y0 + 24)), dyhi = min(R, ph - y0 + 8);
  const int C = dxhi - dxlo + 1, NR = dyhi - dylo + 1;
  const int S = (NR + K - 1) / K;
  // a lane may visit several columns of one strip (when C > NT): its keys order them by row, then
  // by the lane's column step c / NT, so that ties resolve in raster order and decode exactly
  constexpr bool SHIFTED = rmax_of(RAWS) == 32;
  constexpr bool MULTI = 2 * rmax_of(RAWS) + 1 > NT;  // can a lane visit two columns of a row?
  // are a lane's passes more than a wave's width plus a window row apart (see the reduction)?
  constexpr bool LANE_ORDER = !MULTI && 2 * rmax_of(RAWS) + 1 <= NT - 63;
  const int TQ = MULTI ? (C + NT - 1) / NT : 1;
  const uint8_t *g = a.ref_g[r];
  const int tid = threadIdx.x;

  // stage the window's rows; the window's first column sits `wxo` bytes past the dword boundary `al`
  const int wx = x0 + dxlo + PAD, wy = y0 + dylo + PAD;
  const int al = wx & ~3, wxo = wx - al;
  if constexpr (SHIFTED) {
    // item (row, d): the aligned plane dwords d and d + 1 give the row's dwords 4d .. 4d + 3
    constexpr int ND = RAWS / 4;
    for (int e = tid; e < rows_lds * ND; e += NT) {
      const int rr = e / ND, d = e - rr * ND;
      const uint8_t *row = g + (size_t)min(wy + rr, sh - 1) * sw;
      const uint32_t lo = ld_u32(row + min(al + 4 * d, sw - 4)), hi = ld_u32(row + min(al + 4 * d + 4, sw - 4));
      u32x4 v;
      v.x = lo;
      v.y = __builtin_amdgcn_alignbyte(hi, lo, 1);
      v.z = __builtin_amdgcn_alignbyte(hi, lo, 2);
      v.w = __builtin_amdgcn_alignbyte(hi, lo, 3);
      *reinterpret_cast<u32x4 *>(raw + (size_t)rr * RAWS + 4 * d) = v;
    }
  } else {
    for (int e = tid; e < rows_lds * RAWS; e += NT) {
      const int rr = e / RAWS, k = e - rr * RAWS;
      const int gy = min(wy + rr, sh - 1);
      raw[e] = ld_u32(g + (size_t)gy * sw + min(al + 4 * k, sw - 4));
    }
  }
  // the macroblock's samples, uniform across the workgroup
  // (scalar loads: the picture is read-only while the kernel runs)
  uint32_t src[64];
  const CU32 *sp = (const CU32 *)(a.src_y + (size_t)y0 * pw + x0);
#pragma unroll
  for (int i = 0; i < 64; i++) src[i] = sp[(i >> 2) * (pw >> 2) + (i & 3)];
  __syncthreads();

  uint32_t run[IA_NPART];
#pragma unroll
  for (int p = 0; p < IA_NPART; p++) run[p] = 0xffffffffu;

  for (int q = tid; q < S * C; q += NT) {
    const int s = q / C, c = q - s * C;
    const int b0 = wxo + c;                 // byte of the candidate's first sample in a staged row
    const uint32_t sh8 = (uint32_t)(b0 & 3);
    const uint32_t *col = raw + (SHIFTED ? b0 : b0 >> 2) + s * K * RAWS;
    const int i0 = s * K;
    uint32_t acc[K][16];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int b = 0; b < 16; b++) acc[k][b] = (uint32_t)((i0 + k) * TQ + (MULTI ? c / NT : 0));
#pragma unroll
    for (int j = 0; j < K + 15; j++) {
      const uint32_t *rp = col + j * RAWS;
      uint32_t r0, r1, r2, r3;
      if constexpr (SHIFTED) {
        r0 = rp[0], r1 = rp[4], r2 = rp[8], r3 = rp[12];
      } else {
        const uint32_t w0 = rp[0], w1 = rp[1], w2 = rp[2], w3 = rp[3], w4 = rp[4];
        r0 = __builtin_amdgcn_alignbyte(w1, w0, sh8), r1 = __builtin_amdgcn_alignbyte(w2, w1, sh8);
        r2 = __builtin_amdgcn_alignbyte(w3, w2, sh8), r3 = __builtin_amdgcn_alignbyte(w4, w3, sh8);
      }
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int rr = j - k;
        if (rr >= 0 && rr < 16) {
          const int by = rr >> 2;
          acc[k][by * 4 + 0] = __builtin_amdgcn_sad_hi_u8(src[rr * 4 + 0], r0, acc[k][by * 4 + 0]);
          acc[k][by * 4 + 1] = __builtin_amdgcn_sad_hi_u8(src[rr * 4 + 1], r1, acc[k][by * 4 + 1]);
          acc[k][by * 4 + 2] = __builtin_amdgcn_sad_hi_u8(src[rr * 4 + 2], r2, acc[k][by * 4 + 2]);
          acc[k][by * 4 + 3] = __builtin_amdgcn_sad_hi_u8(src[rr * 4 + 3], r3, acc[k][by * 4 + 3]);
        }
      }
    }
    if (i0 + K <= NR) {
      // every candidate of the strip lies in the window: two candidates per v_min3
#pragma unroll
      for (int k = 0; k < K; k += 2) {
        uint32_t ka[IA_NPART], kb[IA_NPART];
        partition_keys(acc[k], ka);
        if (k + 1 < K) {
          partition_keys(acc[k + 1], kb);
#pragma unroll
          for (int p = 0; p < IA_NPART; p++) run[p] = min3u(run[p], ka[p], kb[p]);
        } else {
#pragma unroll
          for (int p = 0; p < IA_NPART; p++) run[p] = min(run[p], ka[p]);
        }
      }
    } else {
      const int nvalid = NR - i0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (k < nvalid) {
          uint32_t ka[IA_NPART];
          partition_keys(acc[k], ka);
#pragma unroll
          for (int p = 0; p < IA_NPART; p++) run[p] = min(run[p], ka[p]);
        }
      }
    }
  }

  // reduce over the wave, then over the workgroup as (SAD << 15) | raster index
  __shared__ uint32_t red[NT / 64][IA_NPART];
  const int lane = tid & 63, wave = tid >> 6;
  if constexpr (LANE_ORDER) {
    // The low half of a running minimum is the absolute row times the blocks summed, so keys compare
    // across lanes up to the column; lanes of one wave that met the same row did so in the same pass
    // (two passes of a lane are NT > C + 62 positions apart), where a lower lane holds a lower column.
    // So the wave's least key, taken in its lowest lane, is the wave's winner; it is decoded once,
    // from uniform values. The DPP steps run over all partitions at once.
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);  // the decode then runs on the scalar unit
    uint32_t m[IA_NPART];
#pragma unroll
    for (int p = 0; p < IA_NPART; p++) m[p] = run[p];
    dpp_min_all<0x111, 0xf, 0xf>(m);
    dpp_min_all<0x112, 0xf, 0xf>(m);
    dpp_min_all<0x114, 0xf, 0xe>(m);
    dpp_min_all<0x118, 0xf, 0xc>(m);
    dpp_min_all<0x142, 0xa, 0xf>(m);
    dpp_min_all<0x143, 0xc, 0xf>(m);
#pragma unroll
    for (int p = 0; p < IA_NPART; p++) {
      const int lg = p == 0 ? 4 : p < 5 ? 3 : p < 9 ? 2 : p < 25 ? 1 : 0;  // log2 of the 4x4 blocks summed
      const uint32_t wm = (uint32_t)__builtin_amdgcn_readlane((int)m[p], 63);
      uint32_t gk = 0xffffffffu;
      if (wm != 0xffffffffu) {
        const int tl = wave_u * 64 + __ffsll((unsigned long long)__ballot(run[p] == wm)) - 1;
        const int i = (int)((wm & 0xffffu) >> lg), s = i / K;
        const int t = s * C > tl ? (s * C - tl + NT - 1) / NT : 0;
        gk = ((wm >> 16) << 15) | (uint32_t)(i * C + tl + t * NT - s * C);
      }
      if (lane == 0) red[wave][p] = gk;
    }
  } else {
    // each lane rewrites its minima as (SAD << 15) | raster index first
#pragma unroll
    for (int p = 0; p < IA_NPART; p++) {
      const int lg = p == 0 ? 4 : p < 5 ? 3 : p < 9 ? 2 : p < 25 ? 1 : 0;
      const uint32_t v = run[p];
      uint32_t gk = 0xffffffffu;
      if (v != 0xffffffffu) {
        const uint32_t sad = v >> 16;
        const int o = (int)((v & 0xffffu) >> lg), i = o / TQ;
        const int s = i / K;
        const int t = s * C > tid ? (s * C - tid + NT - 1) / NT : 0;
        const int c = tid + t * NT - s * C + (o - i * TQ) * NT;  // the lane's first column of the strip, stepped
        gk = (sad << 15) | (uint32_t)(i * C + c);
      }
      gk = wave_min(gk);
      if (lane == 0) red[wave][p] = gk;
    }
  }
  __syncthreads();
  if (tid < IA_NPART) {
    uint32_t gk = red[0][tid];
#pragma unroll
    for (int w = 1; w < NT / 64; w++) gk = min(gk, red[w][tid]);
    ia_fsresult res;
    if (gk == 0xffffffffu) {
      res.mv[0] = res.mv[1] = 0;
      res.sad = 0xffffffffu;
    } else {
      const int idx = (int)(gk & 0x7fffu), i = idx / C, c = idx - i * C;
      res.mv[0] = (int16_t)((dxlo + c) * 4);
      res.mv[1] = (int16_t)((dylo + i) * 4);
      res.sad = gk >> 15;
    }
    a.out[((size_t)r * a.mb_w * a.mb_h + mb) * IA_NPART + tid] = res;
  }
}

// Block (x, y) searches a macroblock against reference y. Workgroups are dispatched to the 8 XCDs in
// turn and each XCD has its own L2, so block x takes macroblock (x % 8) * (n / 8) + x / 8: each XCD
// walks one band of consecutive macroblocks, whose overlapping windows then meet in its L2.
template <int K, int NT, int RAWS, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void fullsearch_kernel(FsArgs a,
                                                                                                 int rows_lds) {
  extern __shared__ __align__(16) uint32_t raw[];
  const int x = blockIdx.x, band = gridDim.x / NXCD;
  const int mb = x < band * NXCD ? (x % NXCD) * band + x / NXCD : x;
  search_one<K, NT, RAWS>(a, rows_lds, mb, blockIdx.y, raw);
}

// Many pictures at once. Workgroup b belongs to member k of the launch (offs[k] <= b < offs[k+1];
// each member's count a multiple of 8, so a block's XCD is that of its index within the member),
// which searches picture sidx[k] with args[sidx[k]]: x = b - offs[k] = ref * nmbp + i, nmbp = the
// macroblocks rounded up to a multiple of 8, i mapped to a macroblock by XCD band as above.
// Workgroups past the picture's macroblocks or its references of this step return at once.
template <int K, int NT, int RAWS, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void fullsearch_batch_kernel(
    const FsArgs *args, const int *offs, const int *sidx, int nmem, int rows_lds) {
  extern __shared__ __align__(16) uint32_t raw[];
  const int b = blockIdx.x;
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const FsArgs &a = args[sidx[lo]];
  const int nmb = a.mb_w * a.mb_h, nmbp = (nmb + NXCD - 1) / NXCD * NXCD;
  const int x = b - offs[lo], r = x / nmbp, i = x - r * nmbp;
  if (r >= a.nref) return;
  const int band = nmbp / NXCD, mb = (i % NXCD) * band + i / NXCD;
  if (mb >= nmb) return;
  search_one<K, NT, RAWS>(a, rows_lds, mb, r, raw);
}

// The fallback list of a batch: entries list[i] = picture << 20 | mb << 2 | ref (i < *count) with the
// picture's arguments args[picture]; the blocks walk the list grid-stride.
template <int K, int NT, int RAWS, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void fullsearch_list_batch_kernel(
    const FsArgs *args, int rows_lds, const int *list, const int *count) {
  extern __shared__ __align__(16) uint32_t raw[];
  const int n = *count;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int e = list[i];
    search_one<K, NT, RAWS>(args[e >> 20], rows_lds, (e >> 2) & 0x3ffff, e & 3, raw);
    __syncthreads();  // the next entry restages the window and the reduction scratch
  }
}

// Block i searches entry list[i] = mb * 4 + ref when i < *count and returns at once otherwise (the
// grid covers every possible entry: the count is known on the device only). Block 0 clears *reset.
template <int K, int NT, int RAWS, int WPE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void fullsearch_list_kernel(
    FsArgs a, int rows_lds, const int *list, const int *count, int *reset) {
  extern __shared__ __align__(16) uint32_t raw[];
  if (blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
  if ((int)blockIdx.x >= *count) return;
  const int e = list[blockIdx.x];
  search_one<K, NT, RAWS>(a, rows_lds, e >> 2, e & 3, raw);
}


struct Variant {
  const char *name;
  int k, nt;
  void (*fn32)(FsArgs, int);  // ranges up to 32
  void (*fn64)(FsArgs, int);  // ranges up to 64
  void (*bfn32)(const FsArgs *, const int *, const int *, int, int);  // many pictures, ranges up to 32
  void (*bfn64)(const FsArgs *, const int *, const int *, int, int);
};

#define IA_FS_VARIANT(K, NT, W)                                                                          \
  {"k" #K "_t" #NT "_w" #W, K, NT, fullsearch_kernel<K, NT, RW32, W>, fullsearch_kernel<K, NT, RW64, W>, \
   fullsearch_batch_kernel<K, NT, RW32, W>, fullsearch_batch_kernel<K, NT, RW64, W>}
const Variant VARIANTS[] = {
    IA_FS_VARIANT(4, 256, 3), IA_FS_VARIANT(6, 256, 3), IA_FS_VARIANT(8, 256, 2), IA_FS_VARIANT(4, 128, 3),
    IA_FS_VARIANT(6, 128, 3), IA_FS_VARIANT(6, 256, 2), IA_FS_VARIANT(5, 128, 3), IA_FS_VARIANT(5, 256, 3),
};
#undef IA_FS_VARIANT
constexpr int NVARIANTS = sizeof(VARIANTS) / sizeof(VARIANTS[0]);
constexpr int DEFAULT_VARIANT = 6;  // k5_t128_w3: fastest on MI355X at 1080p +-32 (measured; 65 rows = 13 strips)
// the list kernel's tile: a fallback list is short (a few percent of the (mb, ref) pairs), so wider
// workgroups than the whole-picture search's shorten the one round it takes (fastest measured at
// 1080p +-32, 3 references: 1.3% ahead of K=6 with 256 threads)
constexpr int LIST_K = 5, LIST_NT = 192, LIST_W = 3;

}  // namespace

int fullsearch_variants() { return NVARIANTS; }
const char *fullsearch_variant_name(int v) { return v >= 0 && v < NVARIANTS ? VARIANTS[v].name : nullptr; }

hipError_t launch_fullsearch(const FsArgs &a, int variant, hipStream_t stream) {
  if (variant < 0 || variant >= NVARIANTS) variant = DEFAULT_VARIANT;
  if (a.range < 1 || a.range > RMAX || a.nref < 1 || a.nref > 16) return hipErrorInvalidValue;
  const Variant &v = VARIANTS[variant];
  const bool small = a.range <= 32;
  const int raws = small ? RW32 : RW64;
  const int rows_lds = rows_for(v.k, a.range);
  const size_t shmem = (size_t)rows_lds * raws * sizeof(uint32_t);
  dim3 grid(a.mb_w * a.mb_h, a.nref);
  hipLaunchKernelGGL(small ? v.fn32 : v.fn64, grid, dim3(v.nt), shmem, stream, a, rows_lds);
  return hipGetLastError();
}

int fullsearch_default_variant() { return DEFAULT_VARIANT; }

hipError_t launch_fullsearch_batch(const FsArgs *args, const int *offs, const int *sidx, int nmem, int nblocks,
                                   int variant, int max_range, hipStream_t stream) {
  if (variant < 0 || variant >= NVARIANTS) variant = DEFAULT_VARIANT;
  if (max_range < 1 || max_range > RMAX || nmem < 1 || nblocks < 1) return hipErrorInvalidValue;
  const Variant &v = VARIANTS[variant];
  const bool small = max_range <= 32;
  const int raws = small ? RW32 : RW64;
  const int rows_lds = rows_for(v.k, max_range);
  const size_t shmem = (size_t)rows_lds * raws * sizeof(uint32_t);
  hipLaunchKernelGGL(small ? v.bfn32 : v.bfn64, dim3(nblocks), dim3(v.nt), shmem, stream, args, offs, sidx, nmem,
                     rows_lds);
  return hipGetLastError();
}

hipError_t launch_fullsearch_list_batch(const FsArgs *args, const int *list, const int *count, int max_range,
                                        int max_entries, hipStream_t stream) {
  if (max_range < 1 || max_range > RMAX || max_entries < 1) return hipErrorInvalidValue;
  const bool small = max_range <= 32;
  const int raws = small ? RW32 : RW64;
  const int rows_lds = rows_for(LIST_K, max_range);
  const size_t shmem = (size_t)rows_lds * raws * sizeof(uint32_t);
  dim3 grid(min(max_entries, 4096));
  hipLaunchKernelGGL(small ? (fullsearch_list_batch_kernel<LIST_K, LIST_NT, RW32, LIST_W>)
                           : (fullsearch_list_batch_kernel<LIST_K, LIST_NT, RW64, LIST_W>),
                     grid, dim3(LIST_NT), shmem, stream, args, rows_lds, list, count);
  return hipGetLastError();
}

hipError_t launch_fullsearch_list(const FsArgs &a, const int *list, const int *count, int *reset,
                                  hipStream_t stream) {
  if (a.range < 1 || a.range > RMAX || a.nref < 1 || a.nref > 4) return hipErrorInvalidValue;
  const bool small = a.range <= 32;
  const int raws = small ? RW32 : RW64;
  const int rows_lds = rows_for(LIST_K, a.range);
  const size_t shmem = (size_t)rows_lds * raws * sizeof(uint32_t);
  dim3 grid(a.mb_w * a.mb_h * a.nref);
  hipLaunchKernelGGL(small ? (fullsearch_list_kernel<LIST_K, LIST_NT, RW32, LIST_W>)
                           : (fullsearch_list_kernel<LIST_K, LIST_NT, RW64, LIST_W>),
                     grid, dim3(LIST_NT), shmem, stream, a, rows_lds, list, count, reset);
  return hipGetLastError();
}

}  // namespace ia
