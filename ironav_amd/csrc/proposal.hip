// proposal.hip — the learned motion proposal on gfx950: cost volume -> BF16 CNN on MFMA -> top-8
// cells per reference -> exact SAD refinement of all 41 partitions -> fallback list.
//
// Persistent: one workgroup of 8 waves per CU holds the CNN weights in LDS; each wave takes one
// macroblock at a time.
//  1. Cost volume: the exact 16x16 SAD at a 9x9 grid of displacements 4 samples apart around each
//     reference's temporal predictor (reference 0's co-located vector, scaled by the reference's
//     distance), clamped so the grid stays within the vectors the macroblock may take.
//  2. Features on the grid (8 channels): log2(1 + SAD) / 16 per reference; one-hot cells of the three
//     temporal neighbours' vectors (co-located, left and above in reference 0); a grid marker.
//  3. Four 3x3 convolutions, 32 channels, ReLU between, BF16 operands and FP32 accumulation with
//     v_mfma_f32_16x16x32_bf16 (implicit GEMM: M = 16 output channels, N = 16 grid cells, K = 32 =
//     one tap's 32 input channels; the first layer packs 4 taps of 8 channels per K step). Outputs:
//     one score map per reference and a predicted log2 SAD per cell.
//  4. The 8 best-scoring cells per reference (ties to the lower cell index), by wave-wide DPP maxima.
//  5. Exact refinement: every integer displacement within +-2 samples of each selected cell
//     (200 per reference), all 41 partitions' SADs with the keyed minima of the exhaustive search
//     (key = SAD << 16 + blocks * order, order = cell rank * 25 + raster offset); the per-lane minima
//     are reduced through a transpose in LDS.
//  6. If the refined 16x16 SAD exceeds fallback_ratio x 2^(predicted log2 SAD at the best cell), the
//     (macroblock, reference) is queued for the exhaustive search, which overwrites its result.
#include "proposal.h"

namespace ia {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// LDS-qualified types: the CNN layer is a function of its own, and generic pointers there would make
// its loads and stores FLAT instructions
typedef __attribute__((address_space(3))) uint16_t LdsU16;
typedef __attribute__((address_space(3))) float LdsF32;
typedef __attribute__((address_space(3))) bf16x8 LdsBf16x8;
typedef __attribute__((address_space(3))) f32x4 LdsF32x4;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x2 LdsU32x2;

constexpr int WAVES = 8;
// The grid in LDS: rows of CNN_GRID cells with no padding between them, a row of zeros above and
// below (a 3x3 tap off the grid's left or right edge reads cell 0, in the zero row above): so 16
// consecutive cells of an MFMA tile lie at 16 consecutive positions, whose 80-byte activation
// cells fall on distinct LDS banks.
constexpr int PCELLS = CNN_GRID * (CNN_GRID + 2);
constexpr int AST = 40;            // bf16 per cell in an activation buffer (32 channels + 8 of padding)
constexpr int WIN = 52;            // staged reference window side: +-18 displacements around the centre + 16
constexpr int WST = 19;            // dwords per staged window row in LDS (padded against bank conflicts)

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ uint32_t dpp_max(uint32_t v) {
  return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, BANK_MASK, false));
}
// three independent wave-wide maxima (the row-shift sequence suits idempotent operations), their DPP
// steps interleaved so that each step's latency is covered by the other two chains
__device__ __forceinline__ void wave_max3(uint32_t &a, uint32_t &b, uint32_t &c) {
#define IA_MAX3_STEP(CTRL, RM)       \
  a = dpp_max<CTRL, RM, 0xf>(a);     \
  b = dpp_max<CTRL, RM, 0xf>(b);     \
  c = dpp_max<CTRL, RM, 0xf>(c);
  IA_MAX3_STEP(0x111, 0xf)
  IA_MAX3_STEP(0x112, 0xf)
  IA_MAX3_STEP(0x114, 0xf)
  IA_MAX3_STEP(0x118, 0xf)
  IA_MAX3_STEP(0x142, 0xa)
  IA_MAX3_STEP(0x143, 0xc)
#undef IA_MAX3_STEP
  a = (uint32_t)__builtin_amdgcn_readlane((int)a, 63);
  b = (uint32_t)__builtin_amdgcn_readlane((int)b, 63);
  c = (uint32_t)__builtin_amdgcn_readlane((int)c, 63);
}

// 41 partition words from 16 raster 4x4 words (as in the exhaustive search)
__device__ __forceinline__ void partition_keys(const uint32_t *b, uint32_t *key) {
#pragma unroll
  for (int b8 = 0; b8 < 4; b8++) {
    const int o = (b8 >> 1) * 8 + (b8 & 1) * 2;
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

// v_min3_u32 written out (the compiler otherwise reduces the pair first, one instruction more)
__device__ __forceinline__ uint32_t min3u(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_min3_u32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// a float's order as an unsigned integer
__device__ __forceinline__ uint32_t ford(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct alignas(16) WaveLds {
  union {
    uint16_t x0[PCELLS][8];          // input features, bf16, zero border
    struct {                         // once the first layer has read the features:
      float out[CNN_CELLS][CNN_OUT]; // the last layer
      int cand[3][CNN_TOPK];         // the selected cells per reference
    };
  };
  union {
    uint16_t act[PCELLS][AST];     // activations, bf16, zero border; each layer overwrites its input
    uint32_t win[3][WIN][WST];     // before and after the CNN: each reference's window, aligned dwords
  };
};

struct PropLds {
  uint16_t w0[12][CNN_CH][8];      // layer 0: K = 12 taps (9 real) x 8 channels
  uint16_t w1[9][CNN_CH][CNN_CH];
  uint16_t w2[9][CNN_CH][CNN_CH];
  uint16_t w3[9][16][CNN_CH];      // layer 3: 16 outputs (CNN_OUT real)
  float b[CNN_LAYERS][CNN_CH];
  WaveLds wv[WAVES];
};

// The 16-byte chunk of weight row r (64 bytes: 32 input channels) that holds input channels
// 8g..8g+7. Sixteen lanes read chunk g of sixteen consecutive rows at once; unswizzled, rows 64 bytes
// apart put those reads on the same LDS banks (8-way conflicts with 128 bytes of banks). Rotating the
// chunk by (r >> 2) + 2 ((r >> 1) & 1) spreads any 8 consecutive rows over 128 bytes and any 16 over
// 256 bytes.
__device__ __forceinline__ int wswz(int r, int g) { return (g + (r >> 2) + 2 * ((r >> 1) & 1)) & 3; }

__device__ __forceinline__ int padpos(int p) { return p + CNN_GRID; }
__device__ __forceinline__ int tapoff(int t) { return (t / 3 - 1) * CNN_GRID + (t % 3 - 1); }
// the position tap t of the cell at position pp (grid column cx) reads: 0 (a zero cell) off the
// grid's left or right edge; rows above and below the grid are zero rows
__device__ __forceinline__ int tappos(int pp, int cx, int t) {
  const int dx = t % 3 - 1;
  return (dx < 0 && cx == 0) || (dx > 0 && cx == CNN_GRID - 1) ? 0 : pp + tapoff(t);
}

// One 3x3 convolution layer on a wave's 9x9 grid. IN0: the first layer (8-channel input); NT: 16-wide
// output tiles; RELU: activation. Reads `in` (stride `ist`), writes bf16 to `outb` or f32 to `outf`.
template <bool IN0, int NT, bool RELU, bool OUTF>
__device__ void conv_layer(int lane, const LdsU16 *in, int ist, const LdsU16 *w, const LdsF32 *bias, LdsU16 *outb,
                           LdsF32 *outf) {
  const int m = lane & 15, g = lane >> 4;
  constexpr int KSTEPS = IN0 ? 3 : 9;
  // the lane's B fragments of every K step and output tile, held in registers for the layer
  bf16x8 bw[KSTEPS][NT];
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ks++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      if (IN0) {
        const int t = ks * 4 + g;  // this lane's tap; taps 9..11 are zero
        bw[ks][nt] = t < 9 ? *(const LdsBf16x8 *)(w + (t * CNN_CH + nt * 16 + m) * 8) : bf16x8{};
      } else {
        const int r = ks * (NT * 16) + nt * 16 + m;  // weight row: tap, output channel
        bw[ks][nt] = *(const LdsBf16x8 *)(w + r * CNN_CH + wswz(r, g) * 8);
      }
    }
  // the whole layer's output stays in registers until every lane has read the input, so the output
  // may overwrite it. Operands: A = the weights (rows = output channels), B = the activations
  // (columns = cells), so D gives each lane 4 consecutive output channels of one cell.
  // the accumulators start at the bias (this lane's 4 output channels of each tile)
  f32x4 bv[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) bv[nt] = *(const LdsF32x4 *)(bias + nt * 16 + 4 * g);
  // K steps outermost: the 6 x NT accumulators are independent, so consecutive MFMAs never wait on
  // each other; each accumulator still sums its K steps in order
  f32x4 acc[6][NT];
  int pp[6], cx[6];
#pragma unroll
  for (int mt = 0; mt < 6; mt++) {
    const int q = min(mt * 16 + m, CNN_CELLS - 1);  // this lane's B column (grid cell) of tile mt
    pp[mt] = padpos(q);
    cx[mt] = q % CNN_GRID;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = bv[nt];
  }
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ks++) {
    bf16x8 x[6];
#pragma unroll
    for (int mt = 0; mt < 6; mt++) {
      if (IN0) {
        const int t = ks * 4 + g;
        x[mt] = t < 9 ? *(const LdsBf16x8 *)(in + tappos(pp[mt], cx[mt], t) * ist) : bf16x8{};
      } else {
        x[mt] = *(const LdsBf16x8 *)(in + tappos(pp[mt], cx[mt], ks) * ist + g * 8);
      }
    }
    // one activation fragment feeds every output tile
#pragma unroll
    for (int mt = 0; mt < 6; mt++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[ks][nt], x[mt], acc[mt][nt], 0, 0, 0);
  }
  wave_sync();
  // D: this lane holds output channels nt*16 + 4g + i, i = 0..3, of cell mt*16 + m
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int c0 = nt * 16 + 4 * g;
#pragma unroll
    for (int mt = 0; mt < 6; mt++) {
      const int q = mt * 16 + m;
      if (q < CNN_CELLS) {
        f32x4 v = acc[mt][nt];
        if (RELU) {
#pragma unroll
          for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i], 0.f);
        }
        if constexpr (!OUTF) {
          const bf16x2 lo = __builtin_convertvector(f32x2{v[0], v[1]}, bf16x2);
          const bf16x2 hi = __builtin_convertvector(f32x2{v[2], v[3]}, bf16x2);
          u32x2 w2;
          w2.x = __builtin_bit_cast(uint32_t, lo);
          w2.y = __builtin_bit_cast(uint32_t, hi);
          *(LdsU32x2 *)(outb + padpos(q) * AST + c0) = w2;
        }
        if (OUTF && c0 < CNN_OUT) *(LdsF32x4 *)(outf + q * CNN_OUT + c0) = v;
      }
    }
  }
}

// stage reference r's window (WIN x WIN samples from plane row wy, column wx) as aligned dwords;
// returns the byte offset of the window's first sample in its first dword
// Load reference r's window (WIN x WIN samples from plane row wy, column wx) as aligned dwords into
// 13 registers per lane: 16 dwords per row (64 samples from the aligned start cover the 52 + 3
// needed), lane taking dword lane % 16 of rows lane / 16 + 4k. Returns the byte offset of the
// window's first sample in its first dword.
__device__ __forceinline__ int load_window(int lane, const uint8_t *g, int sw, int wx, int wy, uint32_t *w) {
  const int al = wx & ~3;
  const uint8_t *p = g + (size_t)(wy + (lane >> 4)) * sw + al + 4 * (lane & 15);
#pragma unroll
  for (int k = 0; k < WIN / 4; k++) {
    w[k] = ld_u32(p);
    p += 4 * sw;
  }
  return wx - al;
}

__device__ __forceinline__ void put_window(int lane, const uint32_t *w, uint32_t (*win)[WST]) {
  uint32_t *q = &win[lane >> 4][lane & 15];
#pragma unroll
  for (int k = 0; k < WIN / 4; k++) q[k * 4 * WST] = w[k];
}

// The window shifted by `sh` bytes (dword q of a row holds the row's bytes 4q + sh .. 4q + sh + 3):
// the cost volume's cells share one byte alignment per reference, so it reads aligned dwords. The
// next dword of a row sits in the next lane of the same 16-lane DPP row (dwords 0..14 are exact;
// the row's last dword is never read).
__device__ __forceinline__ void put_window_shifted(int lane, const uint32_t *w, uint32_t (*win)[WST], uint32_t sh) {
  uint32_t *q = &win[lane >> 4][lane & 15];
#pragma unroll
  for (int k = 0; k < WIN / 4; k++) {
    const uint32_t nx = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[k], 0x101, 0xf, 0xf, false);  // row_shl:1
    q[k * 4 * WST] = __builtin_amdgcn_alignbyte(nx, w[k], sh);
  }
}

// the 4 dwords of 16 window samples at window column x of row y
__device__ __forceinline__ void win_row16(const uint32_t (*win)[WST], int off, int x, int y, uint32_t &r0, uint32_t &r1,
                                          uint32_t &r2, uint32_t &r3) {
  const int b = off + x, q = b >> 2;
  const uint32_t sh = (uint32_t)(b & 3);
  const uint32_t *p = win[y] + q;
  const uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
  r0 = __builtin_amdgcn_alignbyte(w1, w0, sh);
  r1 = __builtin_amdgcn_alignbyte(w2, w1, sh);
  r2 = __builtin_amdgcn_alignbyte(w3, w2, sh);
  r3 = __builtin_amdgcn_alignbyte(w4, w3, sh);
}

// The CNN's weights into the workgroup's LDS (all threads of the workgroup).
__device__ __forceinline__ void load_weights(PropLds &L, const CnnWeights *w, int tid) {
  {
    const CnnWeights &W = *w;
    // layer 0: taps 9..11 zero; layers 1, 2 copied as 16-byte vectors; layer 3: the first 16 outputs
    for (int i = tid; i < 12 * CNN_CH; i += 64 * WAVES) {
      const int t = i / CNN_CH, co = i % CNN_CH;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (t < 9) v = *reinterpret_cast<const uint4 *>(&W.w[0][t][co][0]);
      *reinterpret_cast<uint4 *>(&L.w0[t][co][0]) = v;
    }
    // rows of 32 input channels: chunk g (8 channels) of row r goes to chunk wswz(r, g)
    for (int i = tid; i < 9 * CNN_CH * CNN_CH / 8; i += 64 * WAVES) {
      const int r = i >> 2, d = r * 4 + wswz(r, i & 3);
      reinterpret_cast<uint4 *>(&L.w1[0][0][0])[d] = reinterpret_cast<const uint4 *>(&W.w[1][0][0][0])[i];
      reinterpret_cast<uint4 *>(&L.w2[0][0][0])[d] = reinterpret_cast<const uint4 *>(&W.w[2][0][0][0])[i];
    }
    for (int i = tid; i < 9 * 16 * CNN_CH / 8; i += 64 * WAVES) {
      const int t = i / (16 * CNN_CH / 8), c = i % (16 * CNN_CH / 8), r = t * 16 + (c >> 2);
      reinterpret_cast<uint4 *>(&L.w3[0][0][0])[r * 4 + wswz(r, c & 3)] =
          reinterpret_cast<const uint4 *>(&W.w[3][t][0][0])[c];
    }
    for (int i = tid; i < CNN_LAYERS * CNN_CH; i += 64 * WAVES) (&L.b[0][0])[i] = (&W.b[0][0])[i];
  }
}

// One macroblock of the picture `a` describes, by one wave (its buffers V, the weights in L).
__device__ void propose_mb(PropLds &L, WaveLds &V, const PropArgs &a, int mb, int tid) {
  if (a.nref < 1) return;  // an IDR picture of a batch: nothing to search
  const int nmb = a.mb_w * a.mb_h;
  const int pw = a.mb_w * 16, ph = a.mb_h * 16, sw = pw + 2 * PAD;
  // the lane index afresh in each iteration (an empty asm the compiler cannot see through), so
  // that lane-derived addresses are recomputed here rather than hoisted out of the loop and held
  // in registers for its whole length
  int lane = tid & 63;
  asm volatile("" : "+v"(lane));
  const int mbx = mb % a.mb_w, mby = mb / a.mb_w;
  const int x0 = mbx * 16, y0 = mby * 16;
  const int lox = -(x0 + 24), hix = pw - x0 + 8, loy = -(y0 + 24), hiy = ph - y0 + 8;

  // the macroblock's samples (lane i loads dword i: row i / 4, columns 4 (i % 4)..+3) and reference
  // 0's records of this macroblock, the left and the above one (lane k < 3), all in flight at once
  const uint32_t sv = ld_u32(a.src_y + (size_t)(y0 + (lane >> 2)) * pw + x0 + (lane & 3) * 4);
  int tkind = 0, tmvw = 0;
  if (a.rmbs0 && lane < 3) {
    const int nx = lane == 1 ? mbx - 1 : mbx, ny = lane == 2 ? mby - 1 : mby;
    if (nx >= 0 && ny >= 0) {
      const ia_mbinfo *m = a.rmbs0 + ny * a.mb_w + nx;
      tkind = m->kind;
      tmvw = *reinterpret_cast<const int *>(&m->mv[5][0]);
    }
  }
  // temporal predictors: those records' vectors (quarter samples) when inter coded
  int tmv[3][2], tok[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    tok[k] = __builtin_amdgcn_readlane(tkind, k) == 5;
    const int w = tok[k] ? __builtin_amdgcn_readlane(tmvw, k) : 0;
    tmv[k][0] = (int16_t)(w & 0xffff);
    tmv[k][1] = (int16_t)((uint32_t)w >> 16);
  }
  // grid centres: the co-located vector scaled by the reference's distance, to a multiple of 4
  // samples, clamped so the grid lies within the macroblock's vectors
  int cx[3], cy[3];
  for (int r = 0; r < 3; r++) {
    const int qx = (r + 1) * tmv[0][0], qy = (r + 1) * tmv[0][1];
    const int px = (int)floorf((qx + 8) / 16.0f) * 4, py = (int)floorf((qy + 8) / 16.0f) * 4;
    cx[r] = min(max(px, lox + 16), hix - 16);
    cy[r] = min(max(py, loy + 16), hiy - 16);
  }
  if (lane == 0 && a.centre)
    for (int r = 0; r < 3; r++) {
      a.centre[(mb * 3 + r) * 2] = (int16_t)cx[r];
      a.centre[(mb * 3 + r) * 2 + 1] = (int16_t)cy[r];
    }
  const int nref = a.feat_in ? 0 : a.nref;
  // the reference windows, held in registers until the refinement (the CNN reuses their LDS)
  uint32_t wreg[3][WIN / 4];
  int woff[3] = {0, 0, 0};
#pragma unroll
  for (int r = 0; r < 3; r++)
    if (r < nref) woff[r] = load_window(lane, a.ref_g[r], sw, x0 + cx[r] - 18 + PAD, y0 + cy[r] - 18 + PAD, wreg[r]);
  // the samples as scalars
  uint32_t src[64];
#pragma unroll
  for (int i = 0; i < 64; i++) src[i] = __builtin_amdgcn_readlane(sv, i);
  // for the cost volume, each window shifted to its cells' byte alignment
#pragma unroll
  for (int r = 0; r < 3; r++)
    if (r < nref) put_window_shifted(lane, wreg[r], V.win[r], (uint32_t)((woff[r] + 2) & 3));
  for (int i = lane; i < PCELLS * 8 / 8; i += 64) reinterpret_cast<uint4 *>(&V.x0[0][0])[i] = make_uint4(0, 0, 0, 0);
  wave_sync();

  // features
  if (a.feat_in) {
    for (int i = lane; i < CNN_CELLS * CNN_IN; i += 64)
      V.x0[padpos(i / CNN_IN)][i % CNN_IN] = f2bf(a.feat_in[(size_t)mb * CNN_CELLS * CNN_IN + i]);
  } else {
    // cost volume from the staged windows: the nref * 81 (reference, cell) pairs spread over the lanes
    for (int t = lane; t < nref * CNN_CELLS; t += 64) {
      const int r = t >= 2 * CNN_CELLS ? 2 : t >= CNN_CELLS ? 1 : 0, c = t - r * CNN_CELLS;
      const int off = r == 2 ? woff[2] : r == 1 ? woff[1] : woff[0];
      // the cell's first sample is window byte off + 2 + 4 gx: dword (off + 2) / 4 + gx of the shifted rows
      const int q0 = ((off + 2) >> 2) + c % CNN_GRID, wyp = 4 * (c / CNN_GRID) + 2;
      uint32_t s = 0;
#pragma unroll
      for (int row = 0; row < 16; row++) {
        const uint32_t *wp = &V.win[r][wyp + row][q0];
        const uint32_t r0 = wp[0], r1 = wp[1], r2 = wp[2], r3 = wp[3];
        s = __builtin_amdgcn_sad_u8(src[row * 4 + 0], r0, s);
        s = __builtin_amdgcn_sad_u8(src[row * 4 + 1], r1, s);
        s = __builtin_amdgcn_sad_u8(src[row * 4 + 2], r2, s);
        s = __builtin_amdgcn_sad_u8(src[row * 4 + 3], r3, s);
      }
      const float f = __log2f(1.0f + (float)s) * (1.0f / 16.0f);
      V.x0[padpos(c)][r] = f2bf(f);
      if (a.feat) a.feat[((size_t)mb * CNN_CELLS + c) * CNN_IN + r] = f;
    }
    for (int c = lane; c < CNN_CELLS; c += 64) {
      const int gx = c % CNN_GRID, gy = c / CNN_GRID;
      float fv[CNN_IN] = {0, 0, 0, 0, 0, 0, 1.0f, 0};
      for (int k = 0; k < 3; k++) {
        if (!tok[k]) continue;
        const int kx = (int)floorf((tmv[k][0] - 4 * cx[0] + 8) / 16.0f) + 4;
        const int ky = (int)floorf((tmv[k][1] - 4 * cy[0] + 8) / 16.0f) + 4;
        if (kx == gx && ky == gy) fv[3 + k] = 1.0f;
      }
      for (int ch = 3; ch < CNN_IN; ch++) {
        V.x0[padpos(c)][ch] = f2bf(fv[ch]);
        if (a.feat) a.feat[((size_t)mb * CNN_CELLS + c) * CNN_IN + ch] = fv[ch];
      }
      if (a.feat)
        for (int r = a.nref; r < 3; r++) a.feat[((size_t)mb * CNN_CELLS + c) * CNN_IN + r] = 0.f;
    }
  }
  wave_sync();
  if (a.features_only) return;

  // the CNN (the windows' space now holds the activations, borders zero)
  for (int i = lane; i < PCELLS * AST / 8; i += 64) reinterpret_cast<uint4 *>(&V.act[0][0])[i] = make_uint4(0, 0, 0, 0);
  wave_sync();
  conv_layer<true, 2, true, false>(lane, (const LdsU16 *)&V.x0[0][0], 8, (const LdsU16 *)&L.w0[0][0][0],
                                   (const LdsF32 *)L.b[0], (LdsU16 *)&V.act[0][0], nullptr);
  wave_sync();
  conv_layer<false, 2, true, false>(lane, (const LdsU16 *)&V.act[0][0], AST, (const LdsU16 *)&L.w1[0][0][0],
                                    (const LdsF32 *)L.b[1], (LdsU16 *)&V.act[0][0], nullptr);
  wave_sync();
  conv_layer<false, 2, true, false>(lane, (const LdsU16 *)&V.act[0][0], AST, (const LdsU16 *)&L.w2[0][0][0],
                                    (const LdsF32 *)L.b[2], (LdsU16 *)&V.act[0][0], nullptr);
  wave_sync();
  conv_layer<false, 1, false, true>(lane, (const LdsU16 *)&V.act[0][0], AST, (const LdsU16 *)&L.w3[0][0][0],
                                    (const LdsF32 *)L.b[3], nullptr, (LdsF32 *)&V.out[0][0]);
  wave_sync();
  if (a.cnn_out)
    for (int i = lane; i < CNN_CELLS * CNN_OUT; i += 64) a.cnn_out[(size_t)mb * CNN_CELLS * CNN_OUT + i] = (&V.out[0][0])[i];
  if (a.feat_in) return;

  // the windows again, for the refinement (the activations are done with)
#pragma unroll
  for (int r = 0; r < 3; r++)
    if (r < nref) put_window(lane, wreg[r], V.win[r]);
  wave_sync();

  // the 8 best cells of each reference's score map (ties to the lower cell index), the three
  // references' selections interleaved; lane k < 8 stores the k-th
  {
    uint32_t k0[3], k1[3];
    int mine[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      k0[r] = (ford(V.out[lane][r]) & ~0x7fu) | (uint32_t)(127 - lane);
      k1[r] = lane + 64 < CNN_CELLS ? (ford(V.out[lane + 64][r]) & ~0x7fu) | (uint32_t)(127 - lane - 64) : 0u;
      mine[r] = 0;
    }
    // the three references' selections side by side (those beyond nref are not stored)
#pragma unroll
    for (int k = 0; k < CNN_TOPK; k++) {
      uint32_t m0 = max(k0[0], k1[0]), m1 = max(k0[1], k1[1]), m2 = max(k0[2], k1[2]);
      wave_max3(m0, m1, m2);
      const uint32_t mm[3] = {m0, m1, m2};
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const int cell = 127 - (int)(mm[r] & 0x7f);
        if (cell == lane) k0[r] = 0;
        if (cell == lane + 64) k1[r] = 0;
        if (lane == k) mine[r] = cell;
      }
    }
    wave_sync();
#pragma unroll
    for (int r = 0; r < 3; r++)
      if (r < nref && lane < CNN_TOPK) V.cand[r][lane] = mine[r];
    wave_sync();
  }

  // exact refinement of all 41 partitions around the selected cells, +-2 samples. Lanes 21r..21r+20
  // take reference r; its 40 strips (cell rank k, column x of the cell's 5x5 block) go two per
  // lane, and a strip walks its column's 5 displacements (each staged row read once feeds every
  // candidate that uses it). Lane 63 idles.
  uint32_t run[IA_NPART];
#pragma unroll
  for (int p = 0; p < IA_NPART; p++) run[p] = 0xffffffffu;
  const int rg = lane / 21;
  if (rg < nref) {
    const int off = rg == 2 ? woff[2] : rg == 1 ? woff[1] : woff[0];
    const int crx = rg == 2 ? cx[2] : rg == 1 ? cx[1] : cx[0];
    const int cry = rg == 2 ? cy[2] : rg == 1 ? cy[1] : cy[0];
    for (int strip = lane - 21 * rg; strip < 5 * CNN_TOPK; strip += 21) {
      const int k = strip / 5, x = strip - 5 * k;
      const int cell = V.cand[rg][k];
      const int ox = 4 * (cell % CNN_GRID) + x, oy = 4 * (cell / CNN_GRID);  // window offset of the strip top
      const int dx = crx - 18 + ox, dy0 = cry - 18 + oy;
      uint32_t acc[5][16];
#pragma unroll
      for (int i = 0; i < 5; i++)
#pragma unroll
        for (int b = 0; b < 16; b++) acc[i][b] = (uint32_t)(k * 25 + i * 5 + x);  // order: rank, then raster
#pragma unroll
      for (int y = 0; y < 20; y++) {
        uint32_t r0, r1, r2, r3;
        win_row16(V.win[rg], off, ox, oy + y, r0, r1, r2, r3);
#pragma unroll
        for (int i = 0; i < 5; i++) {
          const int row = y - i;
          if (row >= 0 && row < 16) {
            const int by = row >> 2;
            acc[i][by * 4 + 0] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 0], r0, acc[i][by * 4 + 0]);
            acc[i][by * 4 + 1] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 1], r1, acc[i][by * 4 + 1]);
            acc[i][by * 4 + 2] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 2], r2, acc[i][by * 4 + 2]);
            acc[i][by * 4 + 3] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 3], r3, acc[i][by * 4 + 3]);
          }
        }
      }
      const bool xin = dx >= lox && dx <= hix;
      // candidates 0+1 and 2+3 by v_min3 when both lie within the macroblock's vectors, else one by one
#pragma unroll
      for (int i = 0; i < 5; i += 2) {
        const bool in0 = xin && dy0 + i >= loy && dy0 + i <= hiy;
        const bool in1 = i + 1 < 5 && xin && dy0 + i + 1 >= loy && dy0 + i + 1 <= hiy;
        if (i + 1 < 5 && in0 && in1) {
          uint32_t ka[IA_NPART], kb[IA_NPART];
          partition_keys(acc[i], ka);
          partition_keys(acc[i + 1], kb);
#pragma unroll
          for (int p = 0; p < IA_NPART; p++) run[p] = min3u(run[p], ka[p], kb[p]);
        } else {
#pragma unroll
          for (int j = 0; j < 2; j++) {
            if (i + j < 5 && (j == 0 ? in0 : in1)) {
              uint32_t key[IA_NPART];
              partition_keys(acc[i + j], key);
#pragma unroll
              for (int p = 0; p < IA_NPART; p++) run[p] = min(run[p], key[p]);
            }
          }
        }
      }
    }
  }
  // the minima over each reference's lanes by a transpose through LDS (over the windows, now read):
  // T[lane][p] = run[p]; output o = 41 r + p takes the minimum of column p over rows 21r..21r+20
  wave_sync();
  uint32_t *T = &V.win[0][0][0];
#pragma unroll
  for (int p = 0; p < IA_NPART; p++) T[lane * IA_NPART + p] = run[p];
  wave_sync();
  for (int o = lane; o < nref * IA_NPART; o += 64) {
    const int r = o >= 2 * IA_NPART ? 2 : o >= IA_NPART ? 1 : 0, p = o - IA_NPART * r;
    uint32_t v = 0xffffffffu;
#pragma unroll
    for (int l = 0; l < 21; l++) v = min(v, T[(21 * r + l) * IA_NPART + p]);
    const int lg = p == 0 ? 4 : p < 5 ? 3 : p < 9 ? 2 : p < 25 ? 1 : 0;  // log2 of the 4x4 blocks summed
    ia_fsresult res;
    if (v == 0xffffffffu) {
      res.mv[0] = res.mv[1] = 0;
      res.sad = 0xffffffffu;
    } else {
      const int idx = (int)((v & 0xffffu) >> lg), k = idx / 25, j = idx % 25;
      const int cell = V.cand[r][k];
      const int crx = r == 2 ? cx[2] : r == 1 ? cx[1] : cx[0];
      const int cry = r == 2 ? cy[2] : r == 1 ? cy[1] : cy[0];
      res.mv[0] = (int16_t)((crx + 4 * (cell % CNN_GRID - 4) + (j % 5 - 2)) * 4);
      res.mv[1] = (int16_t)((cry + 4 * (cell / CNN_GRID - 4) + (j / 5 - 2)) * 4);
      res.sad = v >> 16;
    }
    a.out[((size_t)r * nmb + mb) * IA_NPART + p] = res;
    // fallback when the refined 16x16 cost exceeds the ratio times the predicted one
    if (p == 0 && a.fb_list) {
      const float pred = exp2f(V.out[V.cand[r][0]][3]);
      if (res.sad == 0xffffffffu || (float)res.sad > a.fallback_ratio * pred) {
        const int slot = atomicAdd(a.fb_count, 1);
        a.fb_list[slot] = a.fb_tag | (mb * 4 + r);
      }
    }
  }
  wave_sync();  // the next macroblock restages the windows and features
}

// Persistent: each workgroup loads the weights once, then its waves take macroblocks in turn.
__global__ __launch_bounds__(64 * WAVES) void proposal_kernel(PropArgs a) {
  extern __shared__ uint8_t lds_raw[];
  PropLds &L = *reinterpret_cast<PropLds *>(lds_raw);
  const int tid = threadIdx.x, wave = tid >> 6;
  load_weights(L, a.wts, tid);
  __syncthreads();
  WaveLds &V = L.wv[wave];
  const int nmb = a.mb_w * a.mb_h;
  // workgroups go to the 8 XCDs in turn (each with its own L2): renumber them so that those on one
  // XCD take consecutive chunks of macroblocks, whose windows overlap
  const int nxcd = 8, per = gridDim.x / nxcd;
  const int vb = (int)blockIdx.x < per * nxcd ? (int)(blockIdx.x % nxcd) * per + (int)blockIdx.x / nxcd : (int)blockIdx.x;
  for (int mb = vb * WAVES + wave; mb < nmb; mb += gridDim.x * WAVES) propose_mb(L, V, a, mb, tid);
}

// Many pictures in one persistent launch: macroblock g of the launch is macroblock g - offs[k] of
// member k (offs[k] <= g < offs[k+1]), whose picture is args[sidx[k]]. All members share the CNN
// weights of args[sidx[0]].
__global__ __launch_bounds__(64 * WAVES) void proposal_batch_kernel(const PropArgs *args, const int *offs,
                                                                    const int *sidx, int nmem, int total) {
  extern __shared__ uint8_t lds_raw[];
  PropLds &L = *reinterpret_cast<PropLds *>(lds_raw);
  const int tid = threadIdx.x, wave = tid >> 6;
  load_weights(L, args[sidx[0]].wts, tid);
  __syncthreads();
  WaveLds &V = L.wv[wave];
  const int nxcd = 8, per = gridDim.x / nxcd;
  const int vb = (int)blockIdx.x < per * nxcd ? (int)(blockIdx.x % nxcd) * per + (int)blockIdx.x / nxcd : (int)blockIdx.x;
  for (int g = vb * WAVES + wave; g < total; g += gridDim.x * WAVES) {
    const int gu = __builtin_amdgcn_readfirstlane(g);
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offs[mid] <= gu) lo = mid;
      else hi = mid - 1;
    }
    propose_mb(L, V, args[sidx[lo]], gu - offs[lo], tid);
  }
}

}  // namespace

hipError_t launch_proposal_batch(const PropArgs *args, const int *offs, const int *sidx, int nmem, int total,
                                 hipStream_t stream) {
  if (nmem < 1 || total < 1) return hipErrorInvalidValue;
  dim3 grid(min((total + WAVES - 1) / WAVES, 256));
  hipLaunchKernelGGL(proposal_batch_kernel, grid, dim3(64 * WAVES), sizeof(PropLds), stream, args, offs, sidx, nmem,
                     total);
  return hipGetLastError();
}

hipError_t launch_proposal(const PropArgs &a, hipStream_t stream) {
  const int nmb = a.mb_w * a.mb_h;
  if (a.nref < 1 || a.nref > 3) return hipErrorInvalidValue;
  // persistent: one workgroup per CU (its LDS holds the weights), two waves per SIMD walk the macroblocks
  dim3 grid(min((nmb + WAVES - 1) / WAVES, 256));
  const size_t shmem = sizeof(PropLds);
  static_assert(sizeof(WaveLds::win) >= 64 * IA_NPART * sizeof(uint32_t), "the transpose fits over the windows");
  static_assert(sizeof(PropLds) <= 160 * 1024, "the weights and four waves' buffers fit one CU's LDS");
  hipLaunchKernelGGL(proposal_kernel, grid, dim3(64 * WAVES), shmem, stream, a);
  return hipGetLastError();
}

}  // namespace ia
