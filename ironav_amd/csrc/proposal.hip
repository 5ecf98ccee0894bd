// proposal.hip — the learned motion proposal on gfx950: cost volume -> BF16 CNN on MFMA -> top-8
// cells per reference -> exact SAD refinement of all 41 partitions -> fallback list.
//
// One wave per macroblock, four macroblocks per workgroup sharing the CNN weights in LDS.
//  1. Cost volume: the exact 16x16 SAD at a 9x9 grid of displacements 4 samples apart around each
//     reference's temporal predictor (reference 0's co-located vector, scaled by the reference's
//     distance), clamped so the grid stays within the vectors the macroblock may take.
//  2. Features on the grid (8 channels): log2(1 + SAD) / 16 per reference; one-hot cells of the three
//     temporal neighbours' vectors (co-located, left and above in reference 0); a grid marker.
//  3. Four 3x3 convolutions, 32 channels, ReLU between, BF16 operands and FP32 accumulation with
//     v_mfma_f32_16x16x32_bf16 (implicit GEMM: M = 16 grid cells, N = 16 output channels, K = 32 =
//     one tap's 32 input channels; the first layer packs 4 taps of 8 channels per K step). Outputs:
//     one score map per reference and a predicted log2 SAD per cell.
//  4. The 8 best-scoring cells per reference (ties to the lower cell index), by wave-wide DPP maxima.
//  5. Exact refinement: every integer displacement within +-2 samples of each selected cell
//     (200 per reference), all 41 partitions' SADs with the keyed minima of the exhaustive search
//     (key = SAD << 16 + blocks * order, order = cell rank * 25 + raster offset).
//  6. If the refined 16x16 SAD exceeds fallback_ratio x 2^(predicted log2 SAD at the best cell), the
//     (macroblock, reference) is queued for the exhaustive search, which overwrites its result.
#include "proposal.h"

namespace ia {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;
constexpr int PG = CNN_GRID + 2;   // padded grid side (one cell of zeros around)
constexpr int PCELLS = PG * PG;
constexpr int AST = 40;            // bf16 per cell in an activation buffer (32 channels + 8 of padding)

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ uint32_t ld_u32(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }

// one row of 16 reference samples at byte address p (any alignment) as 4 dwords
__device__ __forceinline__ void ld_row16(const uint8_t *p, uint32_t &r0, uint32_t &r1, uint32_t &r2, uint32_t &r3) {
  const uintptr_t a = (uintptr_t)p;
  const uint32_t *q = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
  const uint32_t sh = (uint32_t)(a & 3);
  const uint32_t w0 = q[0], w1 = q[1], w2 = q[2], w3 = q[3], w4 = q[4];
  r0 = __builtin_amdgcn_alignbyte(w1, w0, sh);
  r1 = __builtin_amdgcn_alignbyte(w2, w1, sh);
  r2 = __builtin_amdgcn_alignbyte(w3, w2, sh);
  r3 = __builtin_amdgcn_alignbyte(w4, w3, sh);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v) {
  return min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffff, (int)v, CTRL, ROW_MASK, BANK_MASK, false));
}
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ uint32_t dpp_max(uint32_t v) {
  return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, BANK_MASK, false));
}
// wave-wide minimum / maximum (the row-shift sequence suits idempotent operations)
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
  v = dpp_min<0x111, 0xf, 0xf>(v);
  v = dpp_min<0x112, 0xf, 0xf>(v);
  v = dpp_min<0x114, 0xf, 0xf>(v);
  v = dpp_min<0x118, 0xf, 0xf>(v);
  v = dpp_min<0x142, 0xa, 0xf>(v);
  v = dpp_min<0x143, 0xc, 0xf>(v);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
  v = dpp_max<0x111, 0xf, 0xf>(v);
  v = dpp_max<0x112, 0xf, 0xf>(v);
  v = dpp_max<0x114, 0xf, 0xf>(v);
  v = dpp_max<0x118, 0xf, 0xf>(v);
  v = dpp_max<0x142, 0xa, 0xf>(v);
  v = dpp_max<0x143, 0xc, 0xf>(v);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
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

// a float's order as an unsigned integer
__device__ __forceinline__ uint32_t ford(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct alignas(16) WaveLds {
  uint16_t x0[PCELLS][8];          // input features, bf16, zero border
  uint16_t act[2][PCELLS][AST];    // activations, bf16, zero border
  float out[CNN_CELLS][CNN_OUT];   // last layer
  int cand[3][CNN_TOPK];           // selected cells per reference
  int16_t ctr[3][2];               // grid centre per reference (samples)
};

struct PropLds {
  uint16_t w0[12][CNN_CH][8];      // layer 0: K = 12 taps (9 real) x 8 channels
  uint16_t w1[9][CNN_CH][CNN_CH];
  uint16_t w2[9][CNN_CH][CNN_CH];
  uint16_t w3[9][16][CNN_CH];      // layer 3: 16 outputs (CNN_OUT real)
  float b[CNN_LAYERS][CNN_CH];
  WaveLds wv[WAVES];
};

__device__ __forceinline__ int padpos(int p) { return (p / CNN_GRID + 1) * PG + (p % CNN_GRID) + 1; }
__device__ __forceinline__ int tapoff(int t) { return (t / 3 - 1) * PG + (t % 3 - 1); }

// One 3x3 convolution layer on a wave's 9x9 grid. IN0: the first layer (8-channel input); NT: 16-wide
// output tiles; RELU: activation. Reads `in` (stride `ist`), writes bf16 to `outb` or f32 to `outf`.
template <bool IN0, int NT, bool RELU>
__device__ void conv_layer(const uint16_t *in, int ist, const uint16_t *w, const float *bias, uint16_t *outb,
                           float *outf) {
  const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
  constexpr int KSTEPS = IN0 ? 3 : 9;
  for (int mt = 0; mt < 6; mt++) {
    const int p = min(mt * 16 + m, CNN_CELLS - 1);  // this lane's A row (grid cell)
    const int pp = padpos(p);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KSTEPS; ks++) {
        bf16x8 a, b;
        if (IN0) {
          const int t = ks * 4 + g;  // this lane's tap; taps 9..11 are zero
          if (t < 9) {
            a = *reinterpret_cast<const bf16x8 *>(in + (size_t)(pp + tapoff(t)) * ist);
            b = *reinterpret_cast<const bf16x8 *>(w + ((size_t)t * CNN_CH + nt * 16 + m) * 8);
          } else {
            a = bf16x8{};
            b = bf16x8{};
          }
        } else {
          a = *reinterpret_cast<const bf16x8 *>(in + (size_t)(pp + tapoff(ks)) * ist + g * 8);
          b = *reinterpret_cast<const bf16x8 *>(w + ((size_t)ks * (NT * 16) + nt * 16 + m) * CNN_CH + g * 8);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
      }
      // D: this lane holds output channel nt*16 + m for cells mt*16 + 4g + i
      const int co = nt * 16 + m;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int q = mt * 16 + 4 * g + i;
        if (q < CNN_CELLS) {
          float v = acc[i] + bias[co];
          if (RELU) v = fmaxf(v, 0.f);
          if (outb) outb[(size_t)padpos(q) * AST + co] = f2bf(v);
          if (outf && co < CNN_OUT) outf[q * CNN_OUT + co] = v;
        }
      }
    }
  }
}

__global__ __launch_bounds__(64 * WAVES) void proposal_kernel(PropArgs a) {
  extern __shared__ uint8_t lds_raw[];
  PropLds &L = *reinterpret_cast<PropLds *>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // weights into LDS
  {
    const CnnWeights &W = *a.wts;
    for (int i = tid; i < 12 * CNN_CH * 8; i += 64 * WAVES) {
      const int t = i / (CNN_CH * 8), r = i % (CNN_CH * 8), co = r / 8, ci = r % 8;
      (&L.w0[0][0][0])[i] = t < 9 ? W.w[0][t][co][ci] : 0;
    }
    for (int i = tid; i < 9 * CNN_CH * CNN_CH; i += 64 * WAVES) {
      (&L.w1[0][0][0])[i] = (&W.w[1][0][0][0])[i];
      (&L.w2[0][0][0])[i] = (&W.w[2][0][0][0])[i];
    }
    for (int i = tid; i < 9 * 16 * CNN_CH; i += 64 * WAVES) {
      const int t = i / (16 * CNN_CH), r = i % (16 * CNN_CH), co = r / CNN_CH, ci = r % CNN_CH;
      (&L.w3[0][0][0])[i] = W.w[3][t][co][ci];
    }
    for (int i = tid; i < CNN_LAYERS * CNN_CH; i += 64 * WAVES) (&L.b[0][0])[i] = (&W.b[0][0])[i];
  }
  __syncthreads();
  const int mb = blockIdx.x * WAVES + wave;
  const int nmb = a.mb_w * a.mb_h;
  if (mb >= nmb) return;
  WaveLds &V = L.wv[wave];
  const int mbx = mb % a.mb_w, mby = mb / a.mb_w;
  const int pw = a.mb_w * 16, ph = a.mb_h * 16, sw = pw + 2 * PAD;
  const int x0 = mbx * 16, y0 = mby * 16;
  const int lox = -(x0 + 24), hix = pw - x0 + 8, loy = -(y0 + 24), hiy = ph - y0 + 8;

  // clear the padded buffers
  for (int i = lane; i < PCELLS * 8; i += 64) (&V.x0[0][0])[i] = 0;
  for (int i = lane; i < 2 * PCELLS * AST; i += 64) (&V.act[0][0][0])[i] = 0;

  // the macroblock's samples, uniform
  uint32_t src[64];
  {
    const uint8_t *sp = a.src_y + (size_t)y0 * pw + x0;
#pragma unroll
    for (int i = 0; i < 64; i++) src[i] = __builtin_amdgcn_readfirstlane(ld_u32(sp + (size_t)(i >> 2) * pw + (i & 3) * 4));
  }
  // temporal predictors: reference 0's vectors (quarter samples) at this macroblock, left, above
  int tmv[3][2] = {{0, 0}, {0, 0}, {0, 0}}, tok[3] = {0, 0, 0};
  if (a.rmbs0) {
    const int nx[3] = {mbx, mbx - 1, mbx}, ny[3] = {mby, mby, mby - 1};
    for (int k = 0; k < 3; k++)
      if (nx[k] >= 0 && ny[k] >= 0) {
        const ia_mbinfo &m = a.rmbs0[ny[k] * a.mb_w + nx[k]];
        if (m.kind == 5) {
          tmv[k][0] = m.mv[5][0];
          tmv[k][1] = m.mv[5][1];
          tok[k] = 1;
        }
      }
  }
  // grid centres: the co-located vector scaled by the reference's distance, to a multiple of 4
  // samples, clamped so the grid lies within the macroblock's vectors
  int cx[3], cy[3];
  for (int r = 0; r < 3; r++) {
    int qx = (r + 1) * tmv[0][0], qy = (r + 1) * tmv[0][1];
    int px = (int)floorf((qx + 8) / 16.0f) * 4, py = (int)floorf((qy + 8) / 16.0f) * 4;
    cx[r] = min(max(px, lox + 16), hix - 16);
    cy[r] = min(max(py, loy + 16), hiy - 16);
  }
  if (lane == 0)
    for (int r = 0; r < 3; r++) {
      V.ctr[r][0] = (int16_t)cx[r];
      V.ctr[r][1] = (int16_t)cy[r];
      if (a.centre) {
        a.centre[(mb * 3 + r) * 2] = (int16_t)cx[r];
        a.centre[(mb * 3 + r) * 2 + 1] = (int16_t)cy[r];
      }
    }
  wave_sync();

  // features
  if (a.feat_in) {
    for (int i = lane; i < CNN_CELLS * CNN_IN; i += 64)
      V.x0[padpos(i / CNN_IN)][i % CNN_IN] = f2bf(a.feat_in[(size_t)mb * CNN_CELLS * CNN_IN + i]);
  } else {
    // cost volume: lane handles cells lane and lane + 64 of each reference
    for (int r = 0; r < a.nref; r++) {
      const uint8_t *g = a.ref_g[r];
      for (int c = lane; c < CNN_CELLS; c += 64) {
        const int dx = cx[r] + 4 * (c % CNN_GRID - 4), dy = cy[r] + 4 * (c / CNN_GRID - 4);
        const uint8_t *rp = g + (size_t)(y0 + dy + PAD) * sw + (x0 + dx + PAD);
        uint32_t s = 0;
        for (int row = 0; row < 16; row++) {
          uint32_t r0, r1, r2, r3;
          ld_row16(rp + (size_t)row * sw, r0, r1, r2, r3);
          s = __builtin_amdgcn_sad_u8(src[row * 4 + 0], r0, s);
          s = __builtin_amdgcn_sad_u8(src[row * 4 + 1], r1, s);
          s = __builtin_amdgcn_sad_u8(src[row * 4 + 2], r2, s);
          s = __builtin_amdgcn_sad_u8(src[row * 4 + 3], r3, s);
        }
        const float f = __log2f(1.0f + (float)s) * (1.0f / 16.0f);
        V.x0[padpos(c)][r] = f2bf(f);
        if (a.feat) a.feat[((size_t)mb * CNN_CELLS + c) * CNN_IN + r] = f;
      }
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

  // the CNN
  conv_layer<true, 2, true>(&V.x0[0][0], 8, &L.w0[0][0][0], L.b[0], &V.act[0][0][0], nullptr);
  wave_sync();
  conv_layer<false, 2, true>(&V.act[0][0][0], AST, &L.w1[0][0][0], L.b[1], &V.act[1][0][0], nullptr);
  wave_sync();
  conv_layer<false, 2, true>(&V.act[1][0][0], AST, &L.w2[0][0][0], L.b[2], &V.act[0][0][0], nullptr);
  wave_sync();
  conv_layer<false, 1, false>(&V.act[0][0][0], AST, &L.w3[0][0][0], L.b[3], nullptr, &V.out[0][0]);
  wave_sync();
  if (a.cnn_out)
    for (int i = lane; i < CNN_CELLS * CNN_OUT; i += 64) a.cnn_out[(size_t)mb * CNN_CELLS * CNN_OUT + i] = (&V.out[0][0])[i];
  if (a.feat_in) return;

  // the 8 best cells per reference
  for (int r = 0; r < a.nref; r++) {
    uint32_t k0 = (ford(V.out[lane][r]) & ~0x7fu) | (uint32_t)(127 - lane);
    uint32_t k1 = lane + 64 < CNN_CELLS ? (ford(V.out[lane + 64][r]) & ~0x7fu) | (uint32_t)(127 - lane - 64) : 0u;
    for (int k = 0; k < CNN_TOPK; k++) {
      const uint32_t best = wave_max(max(k0, k1));
      const int cell = 127 - (int)(best & 0x7f);
      if (lane == 0) V.cand[r][k] = cell;
      if (cell == lane) k0 = 0;
      if (cell == lane + 64) k1 = 0;
    }
  }
  wave_sync();

  // exact refinement of all 41 partitions around the selected cells
  for (int r = 0; r < a.nref; r++) {
    const uint8_t *g = a.ref_g[r];
    uint32_t run[IA_NPART];
#pragma unroll
    for (int p = 0; p < IA_NPART; p++) run[p] = 0xffffffffu;
    for (int idx = lane; idx < CNN_TOPK * 25; idx += 64) {
      const int k = idx / 25, j = idx % 25;
      const int cell = V.cand[r][k];
      const int dx = cx[r] + 4 * (cell % CNN_GRID - 4) + (j % 5 - 2);
      const int dy = cy[r] + 4 * (cell / CNN_GRID - 4) + (j / 5 - 2);
      if (dx < lox || dx > hix || dy < loy || dy > hiy) continue;
      const uint8_t *rp = g + (size_t)(y0 + dy + PAD) * sw + (x0 + dx + PAD);
      uint32_t acc[16];
#pragma unroll
      for (int b = 0; b < 16; b++) acc[b] = (uint32_t)idx;
#pragma unroll
      for (int row = 0; row < 16; row++) {
        uint32_t r0, r1, r2, r3;
        ld_row16(rp + (size_t)row * sw, r0, r1, r2, r3);
        const int by = row >> 2;
        acc[by * 4 + 0] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 0], r0, acc[by * 4 + 0]);
        acc[by * 4 + 1] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 1], r1, acc[by * 4 + 1]);
        acc[by * 4 + 2] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 2], r2, acc[by * 4 + 2]);
        acc[by * 4 + 3] = __builtin_amdgcn_sad_hi_u8(src[row * 4 + 3], r3, acc[by * 4 + 3]);
      }
      uint32_t key[IA_NPART];
      partition_keys(acc, key);
#pragma unroll
      for (int p = 0; p < IA_NPART; p++) run[p] = min(run[p], key[p]);
    }
    uint32_t sad16 = 0xffffffffu;
#pragma unroll
    for (int p = 0; p < IA_NPART; p++) {
      const int lg = p == 0 ? 4 : p < 5 ? 3 : p < 9 ? 2 : p < 25 ? 1 : 0;
      const uint32_t v = wave_min(run[p]);
      if (lane == p) {
        ia_fsresult res;
        if (v == 0xffffffffu) {
          res.mv[0] = res.mv[1] = 0;
          res.sad = 0xffffffffu;
        } else {
          const int idx = (int)((v & 0xffffu) >> lg), k = idx / 25, j = idx % 25, cell = V.cand[r][k];
          res.mv[0] = (int16_t)((cx[r] + 4 * (cell % CNN_GRID - 4) + (j % 5 - 2)) * 4);
          res.mv[1] = (int16_t)((cy[r] + 4 * (cell / CNN_GRID - 4) + (j / 5 - 2)) * 4);
          res.sad = v >> 16;
        }
        a.out[((size_t)r * nmb + mb) * IA_NPART + p] = res;
      }
      if (p == 0) sad16 = v == 0xffffffffu ? v : v >> 16;
    }
    // fallback when the refined cost exceeds the ratio times the predicted one
    const float pred = exp2f(V.out[V.cand[r][0]][3]);
    if (lane == 0 && a.fb_list && (sad16 == 0xffffffffu || (float)sad16 > a.fallback_ratio * pred)) {
      const int slot = atomicAdd(a.fb_count, 1);
      a.fb_list[slot] = mb * 4 + r;
    }
  }
}

}  // namespace

hipError_t launch_proposal(const PropArgs &a, hipStream_t stream) {
  const int nmb = a.mb_w * a.mb_h;
  if (a.nref < 1 || a.nref > 3) return hipErrorInvalidValue;
  dim3 grid((nmb + WAVES - 1) / WAVES);
  const size_t shmem = sizeof(PropLds);
  hipLaunchKernelGGL(proposal_kernel, grid, dim3(64 * WAVES), shmem, stream, a);
  return hipGetLastError();
}

}  // namespace ia
