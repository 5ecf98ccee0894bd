// halfpel.hip — the half-sample planes of a reference picture on gfx950.
//
// Builds, from one 8-bit luma plane, the planes `HalfPel::build` builds
// (crates/ironav-codec/src/h264/enc/subpel.rs:57-115):
//   G  the integer samples with the picture's edges repeated PAD samples outwards,
//   B  the horizontal 6-tap half samples (v+16)>>5, columns 2..sw-4 (0 elsewhere),
//   H  the vertical 6-tap half samples, rows 2..sh-4 (0 elsewhere),
//   J  the vertical 6-tap over the unrounded horizontal sums, (v+512)>>10, rows 2..sh-4; the sums are
//      0 outside columns 2..sw-4 as in the reference's scratch buffer.
// HBM-bound: each output byte is written once, the input read about once (through an LDS tile with
// a 2/3-sample halo). A workgroup makes a 64x16 tile of all four planes; a thread makes 4 adjacent
// samples of a row and writes them as one 32-bit store per plane.
#include "common.h"

namespace ia {
namespace {

constexpr int TW = HP_TW, TH = HP_TH;

__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f) {
  return a - 5 * b + 20 * c + 20 * d - 5 * e + f;
}

struct HpLds {
  // G over rows y0-2 .. y0+TH+2 and columns x0-2 .. x0+TW+2 (TH+5 x TW+5), fetched at clamped
  // coordinates: equal to G inside the plane, and only feeding zeroed outputs outside it.
  int16_t gt[TH + 5][TW + 8];
  // unrounded horizontal sums over the same rows, columns x0 .. x0+TW-1
  int32_t b1t[TH + 5][TW];
};

// Tile (bx, by) of the four planes of `luma` (256 threads).
__device__ __forceinline__ void halfpel_tile(HpLds &L, const uint8_t *__restrict__ luma, int pw, int ph,
                                             uint8_t *__restrict__ planes, int bx, int by) {
  const int sw = pw + 2 * PAD, sh = ph + 2 * PAD;
  const size_t n = (size_t)sw * sh;
  const int x0 = bx * TW, y0 = by * TH;
  auto &gt = L.gt;
  auto &b1t = L.b1t;
  const int tid = threadIdx.x;
  for (int i = tid; i < (TH + 5) * (TW + 5); i += 256) {
    const int r = i / (TW + 5), c = i % (TW + 5);
    const int gy = min(max(y0 - 2 + r - PAD, 0), ph - 1);
    const int gx = min(max(x0 - 2 + c - PAD, 0), pw - 1);
    gt[r][c] = luma[(size_t)gy * pw + gx];
  }
  __syncthreads();
  for (int i = tid; i < (TH + 5) * TW; i += 256) {
    const int r = i / TW, c = i % TW;
    const int x = x0 + c;
    int v = 0;
    if (x >= 2 && x < sw - 3)
      v = tap6(gt[r][c], gt[r][c + 1], gt[r][c + 2], gt[r][c + 3], gt[r][c + 4], gt[r][c + 5]);
    b1t[r][c] = v;
  }
  __syncthreads();
  const int ty = tid / (TW / 4), tx = (tid % (TW / 4)) * 4;
  const int y = y0 + ty;
  if (y >= sh || x0 + tx >= sw) return;
  const bool vrow = y >= 2 && y < sh - 3;
  uint32_t og = 0, ob = 0, oh = 0, oj = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = tx + k, x = x0 + c;
    const int g = gt[ty + 2][c + 2];
    int b = 0, h = 0, j = 0;
    if (x >= 2 && x < sw - 3) b = min(max((b1t[ty + 2][c] + 16) >> 5, 0), 255);
    if (vrow) {
      h = tap6(gt[ty][c + 2], gt[ty + 1][c + 2], gt[ty + 2][c + 2], gt[ty + 3][c + 2], gt[ty + 4][c + 2],
               gt[ty + 5][c + 2]);
      h = min(max((h + 16) >> 5, 0), 255);
      j = tap6(b1t[ty][c], b1t[ty + 1][c], b1t[ty + 2][c], b1t[ty + 3][c], b1t[ty + 4][c], b1t[ty + 5][c]);
      j = min(max((j + 512) >> 10, 0), 255);
    }
    og |= (uint32_t)g << (8 * k);
    ob |= (uint32_t)b << (8 * k);
    oh |= (uint32_t)h << (8 * k);
    oj |= (uint32_t)j << (8 * k);
  }
  const size_t o = (size_t)y * sw + x0 + tx;  // sw and x0+tx are multiples of 4
  *reinterpret_cast<uint32_t *>(planes + o) = og;
  *reinterpret_cast<uint32_t *>(planes + n + o) = ob;
  *reinterpret_cast<uint32_t *>(planes + 2 * n + o) = oh;
  *reinterpret_cast<uint32_t *>(planes + 3 * n + o) = oj;
}

__global__ __launch_bounds__(256) void halfpel_kernel(const uint8_t *__restrict__ luma, int pw, int ph,
                                                       uint8_t *__restrict__ planes) {
  __shared__ HpLds lds;
  halfpel_tile(lds, luma, pw, ph, planes, blockIdx.x, blockIdx.y);
}

// Many pictures: block b builds tile b - first of the job whose tiles hold it (jobs by `first`).
__global__ __launch_bounds__(256) void halfpel_batch_kernel(const HpJob *jobs, int njobs) {
  __shared__ HpLds lds;
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
  halfpel_tile(lds, j.luma, j.pw, j.ph, j.planes, t % j.tiles_x, t / j.tiles_x);
}

}  // namespace

hipError_t launch_halfpel_batch(const HpJob *jobs, int njobs, int nblocks, hipStream_t stream) {
  hipLaunchKernelGGL(halfpel_batch_kernel, dim3(nblocks), dim3(256), 0, stream, jobs, njobs);
  return hipGetLastError();
}

hipError_t launch_halfpel(const uint8_t *luma, int pw, int ph, uint8_t *planes, hipStream_t stream) {
  const int sw = pw + 2 * PAD, sh = ph + 2 * PAD;
  dim3 grid((sw + TW - 1) / TW, (sh + TH - 1) / TH);
  hipLaunchKernelGGL(halfpel_kernel, grid, dim3(256), 0, stream, luma, pw, ph, planes);
  return hipGetLastError();
}

}  // namespace ia
